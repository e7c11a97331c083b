// bq_host.cpp - host-side set-up of the Biot time loop (no device work).
//
// These are the reference's one-off set-up computations whose outputs the time
// loop consumes: the PML damping profiles (PPML), the wave speeds they are
// scaled with (CALCUL_VELOCITY) and the source time function (OPEN_SR_FILE).
// They run once per run on the host in binary32, in the reference's expression
// order, so the device sees exactly the coefficients the Fortran host would
// have produced. Built with -ffp-contract=off: every operation rounds as
// written.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <queue>
#include <vector>

#include "bq_internal.h"

namespace bq {

// CALCUL_VELOCITY, calcul_velocity.F90:16-36: eigen-speeds of A^-1 B for the
// 2x2 Biot system, in the no-friction limit.
void wave_speeds(float sig1, float p, float ainv, float mu, float rel1, float rel2,
                 float rel3, float *vpf, float *vps, float *vs) {
  const float ai[2][2] = {{rel1, -rel3}, {-rel3, rel2}};
  const float bm[2][2] = {{sig1, p}, {p, ainv}};
  float tt[2][2];
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) tt[r][c] = ai[r][0] * bm[0][c] + ai[r][1] * bm[1][c];
  const float tr = tt[0][0] + tt[1][1];
  const float delta = tr * tr + 4.0f * (tt[0][1] * tt[1][0] - tt[0][0] * tt[1][1]);
  const float sd = sqrtf(delta);
  *vpf = sqrtf((tr + sd) / 2.0f);
  *vps = sqrtf((tr - sd) / 2.0f);
  *vs = sqrtf(rel1 * mu);
  if (mu == 0.0f) *vs = *vpf;
}

// One damping profile of PPML (ppml.F90:169-207) for a boundary cell of speed
// vel: integer nodes j = 0..T into (a,b,c,d) and half-integer nodes
// j = 0..T+1 into (a2,b2,c2,d2). Node 0 is undamped (a=b=d=1, c=0).
static void damping_profile(int T, float omg, float vel, float sf, float wc, float pw, float dt,
                            float *a, float *b, float *c, float *d, float *a2, float *b2,
                            float *c2, float *d2) {
  for (int j = 0; j <= T; ++j) { a[j] = 1.0f; b[j] = 1.0f; c[j] = 0.0f; }
  for (int j = 0; j <= T + 1; ++j) { a2[j] = 1.0f; b2[j] = 1.0f; c2[j] = 0.0f; }
  auto node = [&](float x, float *pa, float *pb, float *pc) {
    const float om = omg * vel * powf(x, pw);
    const float ap = 1.0f + sf * om * 0.075f / 0.91f;
    const float den = expf(-om / ap - wc * dt);
    *pa = (den + 1.0f) / 2.0f / ap;
    *pb = den;
    *pc = om / (om + wc * ap * dt) * (den - 1.0f);
  };
  for (int j = 1; j <= T; ++j) node((float)j, &a[j], &b[j], &c[j]);
  for (int j = 1; j <= T + 1; ++j) node((float)j - 0.5f, &a2[j], &b2[j], &c2[j]);
  for (int j = 0; j <= T; ++j) d[j] = (c[j] + 2.0f) / (b[j] + 1.0f);
  for (int j = 0; j <= T + 1; ++j) d2[j] = (c2[j] + 2.0f) / (b2[j] + 1.0f);
}

// PPML, ppml.F90:159-306, serial build. Side strips use the S speed of the
// outermost PML column of each row; the bottom strip uses the fast-P speed of
// the last row (ppml.F90:168-169, 215-216, 265-266). KTBO != 0 switches the
// bottom damping off (ppml.F90:261).
void build_pml(const GridDesc &g, const int32_t *jm, const float *tab, float r, float tau,
               float pw, float sf, float wc, PmlHost &out) {
  const int T = g.T, NX = g.NX, NZ = g.NZ;
  out.resize(T, NX, NZ);
  float omg = 0.0f;
  if (T != 0) omg = -tau * logf(r) * g.DT / (float)T / powf((float)T, pw) / g.H;
  auto tabv = [&](int m, int k) { return tab[(size_t)m * 24 + k]; };
  for (int L = 0; L < NZ; ++L) {
    for (int side = 0; side < 2; ++side) {
      const int m = jm[(size_t)L * NX + (side ? NX - 1 : 0)] - 1;
      float vpf, vps, vs;
      wave_speeds(tabv(m, T_SXX1), -tabv(m, T_P1), -tabv(m, T_P3), tabv(m, T_MU),
                  g.H * tabv(m, T_REL1_U) / g.DT, g.H * tabv(m, T_REL2_QX) / g.DT,
                  g.H * tabv(m, T_REL3_U) / g.DT, &vpf, &vps, &vs);
      const size_t o1 = (size_t)L * (T + 1), o2 = (size_t)L * (T + 2);
      const int base = side ? 4 : 0;  // XL arrays 0..3 (a b c d), XH 4..7
      damping_profile(T, omg, vs, sf, wc, pw, g.DT, &out.x1[base + 0][o1], &out.x1[base + 1][o1],
                      &out.x1[base + 2][o1], &out.x1[base + 3][o1], &out.x2[base + 0][o2],
                      &out.x2[base + 1][o2], &out.x2[base + 2][o2], &out.x2[base + 3][o2]);
    }
  }
  if (g.KTBO != 0) omg = 0.0f;
  for (int ix = 0; ix < NX; ++ix) {
    const int m = jm[(size_t)(NZ - 1) * NX + ix] - 1;
    float vpf, vps, vs;
    wave_speeds(tabv(m, T_SZZ2), -tabv(m, T_P2), -tabv(m, T_P3), tabv(m, T_MU),
                g.H * tabv(m, T_REL1_W) / g.DT, g.H * tabv(m, T_REL2_QZ) / g.DT,
                g.H * tabv(m, T_REL3_W) / g.DT, &vpf, &vps, &vs);
    const size_t o1 = (size_t)ix * (T + 1), o2 = (size_t)ix * (T + 2);
    damping_profile(T, omg, vpf, sf, wc, pw, g.DT, &out.z1[0][o1], &out.z1[1][o1],
                    &out.z1[2][o1], &out.z1[3][o1], &out.z2[0][o2], &out.z2[1][o2],
                    &out.z2[2][o2], &out.z2[3][o2]);
  }
}

// --- OPEN_SR_FILE: clamped cubic spline (open_sr_file.F90:121-226) ----------
// The textbook interpolating cubic spline with prescribed end slopes: the
// second derivatives from the tridiagonal system (forward elimination, back
// substitution), evaluation after a bisection search for the bracketing
// interval, as in Press et al., Numerical Recipes, sections 3.3 and 3.4.

static int spl_locate(const std::vector<float> &xx, float x) {
  const int n = (int)xx.size();
  const bool asc = xx[n - 1] >= xx[0];
  int lo = 0, hi = n + 1;
  while (hi - lo > 1) {
    const int mid = (hi + lo) / 2;
    if (asc == (x >= xx[mid - 1])) lo = mid; else hi = mid;
  }
  if (x == xx[0]) return 1;
  if (x == xx[n - 1]) return n - 1;
  return lo;
}

// ------------------------------------------------------------ cell partition

static Split make_split(const Rect *runs, int nruns, const Rect *cells, int ncells, int nr) {
  Split s{};
  s.nr = nr;
  s.nrect = 0;
  s.qcum[0] = 0;
  for (int q = 0; q < nruns; ++q) {
    const Rect &r = runs[q];
    const int ntx = r.x1 > r.x0 ? (r.x1 - r.x0 + 63) / 64 : 0;
    const int nz = r.z1 > r.z0 ? (r.z1 - r.z0 + 4 * nr - 1) / (4 * nr) : 0;
    if (ntx * nz == 0) continue;
    const int k = s.nrect++;
    s.qx0[k] = r.x0; s.qxl[k] = r.x1; s.qntx[k] = ntx; s.qz0[k] = r.z0; s.qz1[k] = r.z1;
    s.qblk[k] = ntx * nz;
    s.qcum[k + 1] = s.qcum[k] + s.qblk[k];
  }
  for (int k = s.nrect; k < kMaxRuns; ++k) s.qcum[k + 1] = s.qcum[k];
  s.cum[0] = 0;
  for (int r = 0; r < kMaxCells; ++r) {
    Rect c{0, 0, 0, 0};
    if (r < ncells && cells[r].x1 > cells[r].x0 && cells[r].z1 > cells[r].z0) c = cells[r];
    s.rx0[r] = c.x0; s.rx1[r] = c.x1; s.rz0[r] = c.z0; s.rz1[r] = c.z1;
    s.cum[r + 1] = s.cum[r] + (c.x1 - c.x0) * (c.z1 - c.z0);
  }
  return s;
}

static long long split_cells(const Split &s) {
  long long n = s.cum[kMaxCells];
  for (int q = 0; q < s.nrect; ++q) n += (long long)(s.qxl[q] - s.qx0[q]) * (s.qz1[q] - s.qz0[q]);
  return n;
}

// Fewest columns a masked quarter column takes; narrower rests go to the runs.
constexpr int kMinMasked = 32;

void fused_rows(Fused &fu, int fr) {
  const int rows = fu.vz1 - fu.vz0;
  auto cdiv = [](int a, int b) { return (a + b - 1) / b; };
  fu.nfz = (fu.nfx > 0 && rows > 0) ? cdiv(rows, fr) : 0;
  fu.nb2 = (fu.h2 && rows > 0) ? cdiv(cdiv(rows, fr / 2), 2) : 0;
  fu.nb4 = (fu.n4 && rows > 0) ? fu.n4 * cdiv(cdiv(rows, fr / 4), 4) : 0;
}

double fused_estimate(const Fused &fu0, int fr) {
  constexpr int kSlots = 512;
  constexpr double kRowUs = 1.9, kRowPerBlockUs = 0.0035;
  Fused fu = fu0;
  fused_rows(fu, fr);
  const int rows = fu.vz1 - fu.vz0;
  // row steps of each block in dispatch order
  std::vector<int> d;
  d.reserve(fused_blocks(fu));
  for (int bz = 0; bz < fu.nfz; ++bz) d.insert(d.end(), fu.nfx, std::min(fr, rows - bz * fr) + 4);
  d.insert(d.end(), fu.nb2, fr / 2 + 4);
  d.insert(d.end(), fu.nb4, fr / 4 + 4);
  // every running block advances at the same rate: keep each block's end in
  // row steps of progress, and the time per row step for the running count
  std::priority_queue<double, std::vector<double>, std::greater<double>> ends;
  double t = 0.0, prog = 0.0;
  size_t next = 0;
  while (next < d.size() || !ends.empty()) {
    while (ends.size() < (size_t)kSlots && next < d.size()) ends.push(prog + d[next++]);
    const double e = ends.top();
    t += (e - prog) * (kRowUs + kRowPerBlockUs * (double)ends.size());
    prog = e;
    while (!ends.empty() && ends.top() <= prog) ends.pop();
  }
  return t;
}

int fused_pick(const Fused &fu) {
  int fr = kTileRows[0];
  double best = fused_estimate(fu, fr);
  for (int c = 1; c < kNumRows; ++c) {
    const double e = fused_estimate(fu, kTileRows[c]);
    if (e < best) {
      best = e;
      fr = kTileRows[c];
    }
  }
  return fr;
}

bool fused_worth(const GridDesc &g, const Fused &fu) {
  Fused f = fu;
  fused_rows(f, 16);
  return fused_blocks(f) > 0 && (long long)g.NX * g.NZ > kCellsMax;
}

// Interior columns start at the first column >= 5 whose address is 128-byte
// aligned, so a wavefront's 64 values of a row are exactly two cache lines.
// Fused tiles cover columns [fx, vxe), vxe <= MX-3, and rows [5, MZ'-3): their
// new stresses reach two rows and columns further, still inside [3, MX-2] x
// [3, MZ'-2], where both half-steps use the interior formulas. Full tiles
// (fw columns) come first, then a column of half tiles and one of quarter
// tiles where the rest of [fx, MX-3) holds them (the last quarter column
// masked to what is left, when that is 32 columns or more). Runs (interior formulas,
// 64-column strips, the last one masked) take the columns [vxe, MX-1) right
// of the tiles (none after a masked column, which ends at MX-3) and the rows
// just above and below the tiles; every other cell - the PML strips, the
// surface and bottom rows, the columns right of the runs - goes to the
// per-cell path. The runs read rows up to z1+2 and columns x0-2 .. x0+64*ntx+1,
// the columns clamped to the grid's last column; with z1 <= MZ'-1, TPML >= 1,
// 4 <= x0 and qxl <= MX-1 the values they use stay inside the grid.
// Edge blocks of the velocity half-step (Split::fold): the last top block
// starts at elast in [max(PX1 - kEdgeCols, 5), MX-2] and the first ends at
// PMLL + kEdgeCols >= 5, so the columns next to every boundary between top
// blocks lie in [4, MX-2], where the velocity formulas are the interior ones
// and no PML memory is touched; the rows must leave 2, 3 and PMLZH-2,
// PMLZH-1 apart from each other. Returns false where the grid does not allow it.
static bool edge_blocks(const GridDesc &g, int zv, Split &sv) {
  const int PX1 = g.PMLXH + 1;
  const int last = std::max(PX1 - kEdgeCols, 5);
  const bool top = zv == 2;  // stress imaging: row 1 from rows 2, 3
  if (top && (last > g.MX - 2 || g.PMLL + kEdgeCols < 5)) return false;
  if (g.PMLZH - 2 < (top ? 4 : 2)) return false;
  sv.fold = 1;
  sv.etop = top ? 1 + (last - g.PMLL + kEdgeCols - 1) / kEdgeCols : 0;
  sv.elast = top ? last : 0;
  sv.ebot = (g.NX + kEdgeCols - 1) / kEdgeCols;
  return true;
}

Plan plan_cells(const GridDesc &g, int off, int fw, int fr, int nr, int zs, int zv, bool fold,
                bool fused) {
  Plan p{};
  const int PX1 = g.PMLXH + 1, ZS1 = g.PMLZH + 1, ZV1 = g.PMLZH;  // exclusive ends
  int fx = 5;
  while ((fx + off) % 32 != 0) ++fx;
  const int ntx = g.MX - 2 >= fx ? (g.MX - 2 - fx + 1) / 64 : 0;
  Fused &fu = p.fu;
  fu.vx0 = fx;
  fu.vz0 = 5;
  fu.vz1 = g.MZ - 3;
  const bool rows = fused && fu.vz1 > fu.vz0 && ntx > 0;
  const int width = rows && g.MX - 3 > fx ? g.MX - 3 - fx : 0;
  fu.nfx = width / fw;
  int rest = width - fw * fu.nfx;
  fu.h2 = rest >= fw / 2;
  rest -= fu.h2 * (fw / 2);
  fu.n4 = 0;
  fu.w4 = fw / 4;
  if (rest >= fw / 4) {
    fu.n4 = 1;
    rest -= fw / 4;
  }
  if (rest >= kMinMasked) {  // a last quarter column, masked to the rest
    fu.n4 += 1;
    fu.w4 = rest;
    rest = 0;
  }
  fu.vxe = fx + fw * fu.nfx + (fw / 2) * fu.h2 + (fw / 4) * (fu.n4 - (fu.n4 > 0)) + (fu.n4 ? fu.w4 : 0);
  fused_rows(fu, fr);
  // the velocity half-step with its edge rows: rows [zt, zb) are left to the
  // per-cell and run rectangles, 2 and 3 (stress imaging) to the top blocks,
  // PMLZH-2 and PMLZH-1 to the bottom blocks
  Split ev{};
  const bool fe = fold && edge_blocks(g, zv, ev);
  const int zt = fe && zv == 2 ? 4 : zv, zb = fe ? g.PMLZH - 2 : ZV1;
  const int zr = std::max(zt, 2);  // first row of the runs
  if (ntx == 0) {  // everything on the per-cell path
    const Rect cs[1] = {{g.PMLL, PX1, zs, ZS1}};
    const Rect cv[1] = {{g.PMLL, PX1, zt, zb}};
    p.ss = make_split(nullptr, 0, cs, 1, nr);
    p.sv = make_split(nullptr, 0, cv, 1, nr);
  } else {
    const int vxe = fu.vxe;  // = fx without fused tiles
    const bool any = vxe > fx;
    // runs end (exclusive): MX-1, or the tiles' end where a masked quarter
    // column took the interior to MX-3 (the two columns left go per cell)
    const int xr = (vxe + off) % 32 == 0 ? g.MX - 1 : vxe;
    p.fused_cells = any ? (long long)(vxe - fx) * (fu.vz1 - fu.vz0) : 0;
    const int zf = any ? fu.vz1 : 5;  // first row below the fused tiles
    const Rect rs[2] = {{vxe, xr, 5, g.MZ - 1}, {fx, vxe, zf, g.MZ - 1}};
    const Rect cs[4] = {{g.PMLL, fx, zs, ZS1}, {xr, PX1, zs, ZS1}, {fx, xr, zs, 5},
                        {fx, xr, g.MZ - 1, ZS1}};
    p.ss = make_split(rs, 2, cs, 4, nr);
    const Rect rv[3] = {{vxe, xr, zr, g.MZ - 1}, {fx, vxe, zr, 5}, {fx, vxe, zf, g.MZ - 1}};
    if (fe) {
      const Rect cv[4] = {{g.PMLL, fx, zt, g.MZ - 1}, {xr, PX1, zt, g.MZ - 1}, {fx, xr, zt, 2},
                          {g.PMLL, PX1, g.MZ - 1, zb}};
      p.sv = make_split(rv, 3, cv, 4, nr);
    } else {
      const Rect cv[4] = {{g.PMLL, fx, zv, ZV1}, {xr, PX1, zv, ZV1}, {fx, xr, zv, 2},
                          {fx, xr, g.MZ - 1, ZV1}};
      p.sv = make_split(rv, 3, cv, 4, nr);
    }
  }
  p.stress_cells = split_cells(p.ss);
  p.velocity_cells = split_cells(p.sv);
  if (fe) {
    p.sv.fold = 1;
    p.sv.etop = ev.etop;
    p.sv.ebot = ev.ebot;
    p.sv.elast = ev.elast;
    p.velocity_cells += (long long)g.NX * (ev.etop ? 4 : 2);  // rows 2, 3 and PMLZH-2, PMLZH-1
  }
  return p;
}

// Two steps per sweep. The interior formulas hold for stresses at columns
// [3, MX-2], rows [3, MZ'-2] and for velocities at columns [4, MX-2], rows
// [2, MZ'-2]. A two-step tile forms level n+1 stresses 6 cells and
// velocities 4 cells around its own cells, so the region is
// [x0, x1) x [9, MZ'-7) with x0 >= 9 and x1 <= MX-7; it reads level-n
// velocities 8 cells around, inside [1, MX] x [1, MZ']. Columns: x0 is the
// first 128-byte aligned column >= max(9, fx), and x1 = x0 + nfx * fw2.
// Outside the region:
//   runs      [fx, xe) above the region (stress rows 5.., velocity rows 2..)
//             and below it (to MZ'-1), and beside it on the 64-column grid
//   per-cell  left strip [PMLL, fx), the surface rows, the bottom rows from
//             MZ'-1, and the right strip from xe (from max(x1, xe) beside
//             the region)
Plan2 plan_cells2(const GridDesc &g, int off, int fw2, int fr2, int nr, int zs, int zv) {
  Plan2 p{};
  const int PX1 = g.PMLXH + 1, ZS1 = g.PMLZH + 1, ZV1 = g.PMLZH;
  int fx = 5;
  while ((fx + off) % 32 != 0) ++fx;
  const int ntx = g.MX - 2 >= fx ? (g.MX - 2 - fx + 1) / 64 : 0;
  const int xe = fx + 64 * ntx;
  // the tiles' first own column on a 128-byte boundary, so that each
  // wavefront's loads of a row are whole cache lines
  int x0 = fx > 9 ? fx : 9;
  while ((x0 + off) % 32 != 0) ++x0;
  Fused2 &f = p.f2;
  f.x0 = x0;
  f.nfx = g.MX - 7 > x0 ? (g.MX - 7 - x0) / fw2 : 0;
  f.x1 = x0 + fw2 * f.nfx;
  f.z0 = 9;
  f.z1 = g.MZ - 7;
  f.fr = fr2;
  f.nfz = (f.nfx > 0 && f.z1 > f.z0 && ntx > 0) ? (f.z1 - f.z0 + fr2 - 1) / fr2 : 0;
  if (f.nfz == 0) {
    f.nfx = 0;
    return p;  // no region: the caller keeps single steps
  }
  p.region_cells = (long long)(f.x1 - f.x0) * (f.z1 - f.z0);
  // beside the region: runs on the 64-column grid from fx ([fx, x0l) on the
  // left, [x1r, xe) on the right), per-cell columns in the gaps [x0l, x0)
  // and [x1, x1r)
  const int x0l = fx + (x0 - fx) / 64 * 64;
  int x1r = fx + (f.x1 - fx + 63) / 64 * 64;
  if (x1r > xe) x1r = xe;
  if (x1r < f.x1) x1r = f.x1;
  const int xr = x1r > xe ? x1r : xe;
  const Rect rs[4] = {{fx, xe, 5, f.z0}, {fx, xe, f.z1, g.MZ - 1}, {fx, x0l, f.z0, f.z1},
                      {x1r, xe, f.z0, f.z1}};
  const Rect cs[8] = {{g.PMLL, fx, zs, ZS1},    {fx, PX1, zs, 5},        {fx, PX1, g.MZ - 1, ZS1},
                      {xe, PX1, 5, f.z0},        {xe, PX1, f.z1, g.MZ - 1}, {xr, PX1, f.z0, f.z1},
                      {f.x1, x1r, f.z0, f.z1},   {x0l, x0, f.z0, f.z1}};
  p.ss = make_split(rs, 4, cs, 8, nr);
  const Rect rv[4] = {{fx, xe, 2, f.z0}, {fx, xe, f.z1, g.MZ - 1}, {fx, x0l, f.z0, f.z1},
                      {x1r, xe, f.z0, f.z1}};
  const Rect cv[8] = {{g.PMLL, fx, zv, ZV1},    {fx, PX1, zv, 2},        {fx, PX1, g.MZ - 1, ZV1},
                      {xe, PX1, 2, f.z0},        {xe, PX1, f.z1, g.MZ - 1}, {xr, PX1, f.z0, f.z1},
                      {f.x1, x1r, f.z0, f.z1},   {x0l, x0, f.z0, f.z1}};
  p.sv = make_split(rv, 4, cv, 8, nr);
  p.stress_cells = split_cells(p.ss);
  p.velocity_cells = split_cells(p.sv);
  return p;
}

// --------------------------------------------- bottom non-reflecting boundary
//
// The coefficient generators NONREF calls compute in binary64 from binary32
// arguments; the products of binary32 arguments they form first (GP, GS of
// HIGCE_FD, OMG*DT of the optimal boundary) round in binary32, as in the
// reference. The nine results a00 .. a22 are returned in a[0..8].

// HIGDON_FD, higdon_fd.F90:1-46: product of two first-order Higdon operators
// with speeds ALFA and BETA (Higdon's absorbing boundary conditions,
// Math. Comp. 47, 1986), discretised with weight WB between time levels.
static void higdon(float WB, float ALFA, float BETA, float DT, float H, double a[9]) {
  const double be1 = 1.0;
  const double be2 = (double)ALFA / (double)BETA;
  const double ni = (double)ALFA * (double)DT / (double)H;
  const double qx = (WB * (be1 + ni) - ni) / (be1 + ni) / (1.0 - WB);
  const double rx = (WB * (be2 + ni) - ni) / (be2 + ni) / (1.0 - WB);
  const double qt = (WB * (be1 + ni) - be1) / (be1 + ni) / (1.0 - WB);
  const double rt = (WB * (be2 + ni) - be1) / (be2 + ni) / (1.0 - WB);
  const double qxt = WB / (WB - 1.0), rxt = WB / (WB - 1.0);
  a[0] = 0.0;
  a[1] = -(qx + rx);
  a[2] = -(qx * rx);
  a[3] = -(qt + rt);
  a[4] = -(qx * rt + qt * rx + qxt + rxt);
  a[5] = -(qx * rxt + rx * qxt);
  a[6] = -(qt * rt);
  a[7] = -(qt * rxt + rt * qxt);
  a[8] = -(qxt * rxt);
}

// HIGCE_FD, higdce_fd.F90:1-42: Higdon operator with speed BETA times a
// first-order Clayton-Engquist one with speed ALFA (the paraxial absorbing
// condition of Clayton and Engquist, BSSA 67, 1977).
static void higdon_ce(float WB, float ALFA, float BETA, float DT, float H, double a[9]) {
  const double gp = (float)(ALFA * DT / H);
  const double gs = (float)(BETA * DT / H);
  const double gs1 = 1.0 + gs;
  const float omw = 1.0f - WB;
  const double h1x = (gs - WB * gs1) / (gs1 * omw);
  const double h1t = (1.0 - WB * gs1) / (gs1 * omw);
  const double hxt = WB / (1.0 - WB);
  const double c1t = 1.0 - gp, cxt = gp;
  a[0] = 0.0;
  a[1] = h1x;
  a[2] = 0.0;
  a[3] = c1t + h1t;
  a[4] = cxt + hxt - h1x * c1t;
  a[5] = -(h1x * cxt);
  a[6] = -(h1t * c1t);
  a[7] = -(hxt * c1t) - h1t * cxt;
  a[8] = -(hxt * cxt);
}

// HIGCE2_FD, higdce2_fd.F90:1-45: Higdon operator with speed VELH times a
// Clayton-Engquist operator with speed VELCE.
static void higdon_ce2(float WB, float VELH, float VELCE, float DT, float H, double a[9]) {
  const double ni = (double)VELH * (double)DT / (double)H;
  const double gce = (double)VELCE * (double)DT / (double)H;
  const double qx = (WB * (1.0 + ni) - ni) / (1.0 + ni) / (1.0 - WB);
  const double rx = (1.0 - gce) / (1.0 + gce);
  const double qt = (WB * (1.0 + ni) - 1.0) / (1.0 + ni) / (1.0 - WB);
  const double rt = -((1.0 - gce) / (1.0 + gce));
  const double qxt = WB / (WB - 1.0), rxt = -1.0;
  a[0] = 0.0;
  a[1] = -(qx + rx);
  a[2] = -(qx * rx);
  a[3] = -(qt + rt);
  a[4] = -(qx * rt + qt * rx + qxt + rxt);
  a[5] = -(qx * rxt + rx * qxt);
  a[6] = -(qt * rt);
  a[7] = -(qt * rxt + rt * qxt);
  a[8] = -(qxt * rxt);
}

// OptAbsBc, OptAbsBc.F90:1-106: the optimally absorbing boundary of Peng and
// Toksoz (JASA 95, 1994): the 3x3 boundary stencil that minimises the
// reflection of plane waves at the two incidence angles gp and gm, for the
// normalised frequency wt = OMG*DT and wavenumber wx = OMG/VEL*H. r0 and r1
// are the two stencils that each annihilate one angle; the result is
// r0 + ratio*r1 with ratio the least-squares weight -<r0,r1>/<r1,r1>.
static void optimal_boundary(double gp, double gm, double wt, double wx, double a[9]) {
  const double tp = wx * std::cos(gp), tm = wx * std::cos(gm);
  const double rz = std::cos(wt), iz = std::sin(wt);
  const double rx = std::cos(tp + tm), ix = std::sin(tp + tm);
  const double ry = -(std::cos(tp) + std::cos(tm)), iy = -(std::sin(tp) + std::sin(tm));
  const double p1 = rx * rz - ix * iz, p2 = rx * iz + rz * ix;
  const double p3 = ry * rz - iy * iz, p4 = ry * iz + rz * iy;
  double inv = (ry + iy + 2.0 * rz) * (p1 + p2 - rz + iz) + (1.0 - rx - ix) * (2.0 + p3 + p4);
  inv = 1.0 / inv;
  double r1[3][3], r0[3][3];
  r1[1][0] = (1.0 - rz) * (p1 + p2 - rz + iz) * inv;
  r1[0][2] = (1.0 - rz) * (1.0 - rx - ix) * inv;
  r1[0][1] = -0.5 - r1[1][0];
  r1[1][1] = 1.0 - 2.0 * r1[0][2];
  r1[0][0] = r1[2][2] = 0.0;
  r1[1][2] = r1[1][0];
  r1[2][0] = r1[0][2];
  r1[2][1] = r1[0][1];
  const double m5 = rx * rz + ix * iz, m6 = rx * iz - rz * ix;
  const double m7 = ry * rz + iy * iz, m8 = ry * iz - rz * iy;
  const double phi = -((rz + iz - m5 + m6) * (ry + iy + 2.0 * rz)) + (1.0 + rz + m7 - m8) * (1.0 - rx - ix);
  const double psi = (1.0 + rz + m7 - m8) * (p1 + p2 - rz + iz) + (rz + iz - m5 + m6) * (2.0 + p3 + p4);
  r0[0][2] = phi * inv;
  r0[1][0] = psi * inv;
  r0[0][0] = 0.0;
  r0[0][1] = 0.5 - r0[1][0];
  r0[1][1] = 1.0 - 2.0 * r0[0][2];
  r0[2][2] = -1.0;
  r0[1][2] = r0[1][0];
  r0[2][0] = r0[0][2];
  r0[2][1] = r0[0][1];
  double dot = 0.0, nrm = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      dot = dot + r0[i][j] * r1[i][j];
      nrm = nrm + r1[i][j] * r1[i][j];
    }
  const double ratio = -(dot / nrm);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[3 * i + j] = r0[i][j] + ratio * r1[i][j];
  a[0] = 0.0;
}

int nonref(int key, float H, float DT, float VEL, float VELS, float VELP, const NrParams &p,
           float ax[8]) {
  double a[9];
  for (int k = 0; k < 8; ++k) ax[k] = 0.0f;
  switch (key) {
    case 0:
      return 0;
    case 1:
      higdon(p.WB, VELP, VELS, DT, H, a);
      break;
    case 2: {  // first-order one-way (transparent) wave equation, upwind: Reynolds'
               // boundary (Geophysics 43, 1978) in its first-order form
      const float c = -(DT * VEL / H);
      ax[2] = 1.0f + c;
      ax[3] = 1.0f - c;
      ax[6] = -1.0f - c;
      ax[7] = c;
      return 0;
    }
    case 3:
    case 9: {  // both keys take the left-boundary angles THPPLE, THSSLE
      const float wx = key == 3 ? p.OMG / VEL * H : p.OMG / VELS * H;
      optimal_boundary((double)p.THPPLE, (double)p.THSSLE, (double)(p.OMG * DT), (double)wx, a);
      break;
    }
    case 4: {
      const float t1 = (DT - H / VEL) / (DT + H / VEL);
      const float t2 = 2.0f * H / VEL / (DT + H / VEL);
      ax[0] = t1;
      ax[2] = t2;
      ax[3] = t2;
      ax[5] = t1;
      ax[6] = -1.0f;
      return 0;
    }
    case 5: {
      const float c = VEL * DT / H;
      ax[0] = -(1.0f - c) / (1.0f + c);
      ax[2] = (1.0f - c) / (1.0f + c);
      ax[3] = 1.0f;
      return 0;
    }
    case 6:
      higdon_ce(p.WB, VELP, VELS, DT, H, a);
      break;
    case 7:
      higdon_ce2(p.WB, VEL, VEL, DT, H, a);
      break;
    case 8:
      higdon_ce2(p.WB, VELS, VELP, DT, H, a);
      break;
    default:
      return -1;
  }
  for (int k = 0; k < 8; ++k) ax[k] = (float)a[k + 1];
  return 0;
}

int bottom_nr(const GridDesc &g, const int32_t *jm, const float *tab, const NrParams &p,
              float abw[8], float abh[8]) {
  const int m = jm[(size_t)g.MZ * g.NX] - 1;  // JM(PMLL, MZ')
  auto tv = [&](int c) { return tab[(size_t)m * T_NTAB + c]; };
  float vpf, vps, vs;
  wave_speeds(tv(T_SXX1), -tv(T_P1), -tv(T_P3), tv(T_MU), g.H * tv(T_REL1_U) / g.DT,
              g.H * tv(T_REL2_QX) / g.DT, g.H * tv(T_REL3_U) / g.DT, &vpf, &vps, &vs);
  if (nonref(g.KTBO, g.H, g.DT, vpf, vs, vpf, p, abw)) return -1;
  return nonref(g.KTBO, g.H, g.DT, vs, vs, vpf, p, abh);
}

}  // namespace bq

extern "C" int bq_source_from_stf(const float *t, const float *v, int nstf, int mt2, float dt,
                                  float h, int key_sour, float qr, float *sourtf) {
  if (!t || !v || !sourtf || nstf < 1 || mt2 < 1) return -1;
  const int n = nstf + 1;
  std::vector<float> x(t, t + nstf), y(v, v + nstf);
  x.push_back((float)mt2 * dt);  // TIME(NSTFS) = MT2*DT
  y.push_back(y[n - 2]);         // STF(NSTFS) = STF(NSTFS-1)
  // spline(x, y, yp1 = 0, ypn = 0, y2)
  std::vector<float> a(n), b(n), c(n), r(n), r0(n), y2(n), gam(n);
  for (int i = 0; i < n - 1; ++i) c[i] = x[i + 1] - x[i];
  for (int i = 0; i < n - 1; ++i) r[i] = 6.0f * ((y[i + 1] - y[i]) / c[i]);
  r0 = r;
  for (int i = 1; i < n - 1; ++i) r[i] = r0[i] - r0[i - 1];
  for (int i = 1; i < n - 1; ++i) a[i] = c[i - 1];
  for (int i = 1; i < n - 1; ++i) b[i] = 2.0f * (c[i] + a[i]);
  b[0] = 1.0f;
  b[n - 1] = 1.0f;
  r[0] = (3.0f / (x[1] - x[0])) * ((y[1] - y[0]) / (x[1] - x[0]) - 0.0f);
  c[0] = 0.5f;
  r[n - 1] = (-3.0f / (x[n - 1] - x[n - 2])) * ((y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]) - 0.0f);
  a[n - 1] = 0.5f;
  // tridag: sub-diagonal a(2:n), diagonal b, super-diagonal c(1:n-1)
  float bet = b[0];
  if (bet == 0.0f) return -2;
  y2[0] = r[0] / bet;
  for (int j = 1; j < n; ++j) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    if (bet == 0.0f) return -2;
    y2[j] = (r[j] - a[j] * y2[j - 1]) / bet;
  }
  for (int j = n - 2; j >= 0; --j) y2[j] = y2[j] - gam[j + 1] * y2[j + 1];
  std::memset(sourtf, 0, sizeof(float) * 2 * (size_t)mt2);
  const float h2 = h * h;
  for (int I = 1; I <= mt2; ++I) {
    const float T = (float)(I - 1) * dt;
    if (T > x[n - 2]) break;  // zero padding after the last sample
    const int klo = std::max(std::min(bq::spl_locate(x, T), n - 1), 1);
    const int khi = klo + 1;
    const float hh = x[khi - 1] - x[klo - 1];
    if (hh == 0.0f) return -3;
    const float aa = (x[khi - 1] - T) / hh;
    const float bb = (T - x[klo - 1]) / hh;
    const float s = aa * y[klo - 1] + bb * y[khi - 1] +
                    ((aa * aa * aa - aa) * y2[klo - 1] + (bb * bb * bb - bb) * y2[khi - 1]) *
                        (hh * hh) / 6.0f;
    float *o = sourtf + (size_t)(I - 1) * 2;
    switch (key_sour) {
      case 0: o[1] = qr * s / h2; o[0] = (1.0f - qr) * s / h2; break;
      case 1: o[1] = 0.0f; o[0] = s / h2; break;
      case 2: o[1] = s / h2; o[0] = qr * s / h2; break;
      case 3: o[1] = s / h2; o[0] = 0.0f; break;
      default: return -4;
    }
  }
  return 0;
}

// NONREF for one boundary, without a handle (host-side set-up helper).
extern "C" int bq_nonref(int key, float h, float dt, float vel, float vels, float velp,
                         const float prm[4], float out[8]) {
  if (!prm || !out) return -1;
  const bq::NrParams p{prm[0], prm[1], prm[2], prm[3]};
  return bq::nonref(key, h, dt, vel, vels, velp, p, out) ? -2 : 0;
}
