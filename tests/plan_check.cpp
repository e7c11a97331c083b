// Exhaustive check of the cell partition (bq::plan_cells) on a list of grids,
// with and without fused tiles: every cell of each half-step is updated by
// exactly one of the fused tiles, the runs and the per-cell path, and every
// kernel's reads stay in the grid.
// Usage: plan_check MX MZ' TPML [MX MZ' TPML ...]; prints "ok <grids>".
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "bq_internal.h"

using namespace bq;

static int fail(const char *what, int mx, int mz, int t, int x, int z) {
  std::printf("FAIL %s grid %d x %d T=%d at x=%d z=%d\n", what, mx, mz, t, x, z);
  return 1;
}

static int check(int MX, int MZ, int T, int fw, int fr, int nr, int zs, int zv, bool fused) {
  GridDesc g{};
  g.MX = MX; g.MZ = MZ; g.T = T; g.PMLL = 1 - T; g.PMLXH = MX + T; g.PMLZH = MZ + T;
  g.NX = MX + 2 * T; g.NZ = MZ + T + 1;
  const int off = T - 1;
  const Plan p = plan_cells(g, off, fw, fr, nr, zs, zv, true, fused);
  if (!fused && (fused_blocks(p.fu) != 0 || p.fused_cells != 0)) return fail("tiles left", MX, MZ, T, 0, 0);
  const int NX = g.NX, NZ = g.NZ;
  std::vector<int> cs((size_t)NX * NZ, 0), cv((size_t)NX * NZ, 0);
  auto at = [&](std::vector<int> &v, int x, int z) -> int & { return v[(size_t)z * NX + (x - g.PMLL)]; };
  auto inside = [&](int x, int z) { return x >= g.PMLL && x <= g.PMLXH && z >= 0 && z <= g.PMLZH; };
  // fused tiles: full tiles, then the column of half tiles and the column of
  // quarter tiles (several to a block, consecutive in z; the last block's may
  // run past the rows and then be empty)
  const Fused &f = p.fu;
  long long nf = 0;
  Fused f2 = f;
  fused_rows(f2, fr);
  if (f2.nfz != f.nfz || f2.nb2 != f.nb2 || f2.nb4 != f.nb4) return fail("fused rows", MX, MZ, T, 0, 0);
  if (f.n4 < 0 || f.n4 > 2 || f.w4 < 1 || f.w4 > fw / 4 || (f.n4 == 0 && f.w4 != fw / 4))
    return fail("quarter columns", MX, MZ, T, f.vxe, 0);
  if (f.vxe != f.vx0 + fw * f.nfx + (fw / 2) * f.h2 + (fw / 4) * (f.n4 - (f.n4 > 0)) + (f.n4 ? f.w4 : 0))
    return fail("fused columns", MX, MZ, T, f.vxe, 0);
  // a tile of wd columns stores [c, c+ws); its threads form stresses at
  // c-2 .. c+wd+1 and read old velocities at c-4 .. c+wd+3, clamped to the
  // grid's last column: the values the stored cells use (stresses at
  // c-2 .. c+ws+1, velocities at c-4 .. c+ws+3) must be unclamped, inside the
  // grid, and of the interior formulas (stresses at columns <= MX-2)
  auto tile = [&](int c, int wd, int ws, int a, int e) -> int {
    if (c < 5 || c + ws > MX - 3 || a < 5 || e > MZ - 3) return fail("fused tile range", MX, MZ, T, c, a);
    if ((c + off) % 32 != 0) return fail("fused alignment", MX, MZ, T, c, a);
    if (!inside(c - 4, a - 4) || !inside(c + ws + 3, e + 3)) return fail("fused reads", MX, MZ, T, c, a);
    if (ws < wd && c + ws + 1 > MX - 2) return fail("masked tile halo", MX, MZ, T, c, a);
    for (int z = a; z < e; ++z)
      for (int x = c; x < c + ws; ++x) { ++at(cs, x, z); ++at(cv, x, z); ++nf; }
    return 0;
  };
  for (int bz = 0; bz < f.nfz; ++bz)
    for (int bx = 0; bx < f.nfx; ++bx) {
      const int c = f.vx0 + bx * fw, a = f.vz0 + bz * fr;
      const int e = std::min(a + fr, f.vz1);
      if (e <= a || c + fw > f.vxe) return fail("empty fused tile", MX, MZ, T, c, a);
      if (tile(c, fw, fw, a, e)) return 1;
    }
  // half column, then the quarter columns
  const int rows = f.vz1 - f.vz0;
  auto cdiv = [](int x, int y) { return (x + y - 1) / y; };
  if (f.nb2 != (f.h2 ? cdiv(cdiv(rows, fr / 2), 2) : 0) ||
      f.nb4 != (f.n4 ? f.n4 * cdiv(cdiv(rows, fr / 4), 4) : 0))
    return fail("narrow blocks", MX, MZ, T, f.vxe, 0);
  const int ncol = f.h2 + f.n4;
  for (int q = 0; q < ncol; ++q) {
    const bool half = q < f.h2;
    const int per = half ? 2 : 4, wd = fw / per, hr = fr / per;
    const int nb = half ? f.nb2 : f.nb4 / f.n4;
    const int c = f.vx0 + fw * f.nfx + (half ? 0 : (fw / 2) * f.h2 + wd * (q - f.h2));
    const int ws = (!half && q == ncol - 1) ? f.w4 : wd;
    for (int b = 0; b < nb; ++b)
      for (int j = 0; j < per; ++j) {
        const int a = f.vz0 + (per * b + j) * hr, e = std::min(a + hr, f.vz1);
        if (e <= a) {
          if (b != nb - 1) return fail("empty narrow tile", MX, MZ, T, c, a);
          continue;
        }
        if (tile(c, wd, ws, a, e)) return 1;
      }
  }
  if (nf != p.fused_cells) return fail("fused count", MX, MZ, T, 0, 0);
  auto runs = [&](const Split &s, std::vector<int> &cnt, int zmin, long long *n) -> int {
    for (int q = 0; q < s.nrect; ++q) {
      // stores [x0, qxl); the lanes that use their neighbours' values reach
      // columns x0-2 .. qxl+1, rows qz0-2 .. qz1+2 (loads past the grid's last
      // column are clamped to it and feed only lanes that do not store)
      const int x0 = s.qx0[q], x1 = s.qxl[q];
      if ((x0 + off) % 32 != 0) return fail("run alignment", MX, MZ, T, x0, s.qz0[q]);
      if (x0 < 4 || x1 > MX - 1 || s.qz0[q] < zmin || s.qz1[q] > MZ - 1)
        return fail("run range", MX, MZ, T, x0, s.qz0[q]);
      if (s.qntx[q] != (x1 - x0 + 63) / 64) return fail("run strips", MX, MZ, T, x0, s.qz0[q]);
      if (!inside(x0 - 2, s.qz0[q] - 2) || !inside(x1 + 1, std::min(s.qz1[q] + 2, g.PMLZH)))
        return fail("run reads", MX, MZ, T, x0, s.qz0[q]);
      const int nz = (s.qz1[q] - s.qz0[q] + 4 * s.nr - 1) / (4 * s.nr);
      if (s.qblk[q] != s.qntx[q] * nz || s.qcum[q + 1] != s.qcum[q] + s.qblk[q])
        return fail("run blocks", MX, MZ, T, x0, s.qz0[q]);
      for (int z = s.qz0[q]; z < s.qz1[q]; ++z)
        for (int x = x0; x < x1; ++x) { ++at(cnt, x, z); ++*n; }
    }
    for (int r = 0; r < kMaxCells; ++r)
      for (int z = s.rz0[r]; z < s.rz1[r]; ++z)
        for (int x = s.rx0[r]; x < s.rx1[r]; ++x) {
          if (!inside(x, z)) return fail("cell outside grid", MX, MZ, T, x, z);
          ++at(cnt, x, z);
          ++*n;
        }
    return 0;
  };
  long long ns = 0, nv = 0;
  if (runs(p.ss, cs, 5, &ns) || runs(p.sv, cv, 2, &nv)) return 1;
  // edge blocks of the velocity half-step: top blocks over rows 2, 3 (stress
  // imaging), bottom blocks over rows PMLZH-2, PMLZH-1, all columns; the
  // columns each top block recomputes for its neighbours (one each side of
  // every boundary between blocks) must use the interior formulas, [4, MX-2]
  const Split &sv = p.sv;
  if (p.ss.fold || p.ss.etop || p.ss.ebot) return fail("stress split with edge blocks", MX, MZ, T, 0, 0);
  if (sv.fold) {
    const int PX1 = g.PMLXH + 1;
    if ((sv.etop > 0) != (zv == 2)) return fail("top blocks", MX, MZ, T, 0, 0);
    int x = g.PMLL;
    for (int b = 0; b < sv.etop; ++b) {
      const int c = b == sv.etop - 1 ? sv.elast : g.PMLL + kEdgeCols * b;
      const int ce = b == sv.etop - 1 ? PX1 : std::min(g.PMLL + kEdgeCols * (b + 1), sv.elast);
      if (c != x || ce <= c || ce - c > kEdgeCols) return fail("top block columns", MX, MZ, T, c, 2);
      if (b > 0 && (c - 1 < 4 || c > MX - 2)) return fail("top block halo", MX, MZ, T, c, 2);
      for (int z = 2; z <= 3; ++z)
        for (int xx = c; xx < ce; ++xx) { ++at(cv, xx, z); ++nv; }
      x = ce;
    }
    if (sv.etop && x != PX1) return fail("top blocks end", MX, MZ, T, x, 2);
    if (sv.ebot != (g.NX + kEdgeCols - 1) / kEdgeCols || g.PMLZH - 2 < (sv.etop ? 4 : 2))
      return fail("bottom blocks", MX, MZ, T, 0, g.PMLZH);
    for (int z = g.PMLZH - 2; z <= g.PMLZH - 1; ++z)
      for (int xx = g.PMLL; xx < PX1; ++xx) { ++at(cv, xx, z); ++nv; }
  } else if (sv.etop || sv.ebot) {
    return fail("edge blocks without fold", MX, MZ, T, 0, 0);
  }
  if (ns != p.stress_cells || nv != p.velocity_cells) return fail("cell counts", MX, MZ, T, 0, 0);
  for (int z = 0; z <= g.PMLZH; ++z)
    for (int x = g.PMLL; x <= g.PMLXH; ++x) {
      const int want_s = z >= zs ? 1 : 0, want_v = (z >= zv && z <= g.PMLZH - 1) ? 1 : 0;
      if (at(cs, x, z) != want_s) return fail("stress coverage", MX, MZ, T, x, z);
      if (at(cv, x, z) != want_v) return fail("velocity coverage", MX, MZ, T, x, z);
    }
  return 0;
}

// The two-step partition: every cell of each half-step is either in the
// region or covered once by the runs and the per-cell path; the region's
// halos (level n+1 stresses 6 cells, velocities 4 cells around its own cells)
// use only interior-formula cells; its reads (8 cells around) stay in the grid.
static int check2(int MX, int MZ, int T, int fw, int fr, int nr, int zs, int zv) {
  GridDesc g{};
  g.MX = MX; g.MZ = MZ; g.T = T; g.PMLL = 1 - T; g.PMLXH = MX + T; g.PMLZH = MZ + T;
  g.NX = MX + 2 * T; g.NZ = MZ + T + 1;
  const int off = T - 1;
  const Plan2 p = plan_cells2(g, off, fw, fr, nr, zs, zv);
  const Fused2 &f = p.f2;
  if (f.nfx == 0) return 0;
  const int NX = g.NX, NZ = g.NZ;
  std::vector<int> cs((size_t)NX * NZ, 0), cv((size_t)NX * NZ, 0);
  auto at = [&](std::vector<int> &v, int x, int z) -> int & { return v[(size_t)z * NX + (x - g.PMLL)]; };
  auto inside = [&](int x, int z) { return x >= g.PMLL && x <= g.PMLXH && z >= 0 && z <= g.PMLZH; };
  if ((f.x0 + off) % 32 != 0) return fail("region alignment", MX, MZ, T, f.x0, f.z0);
  if (f.x1 != f.x0 + fw * f.nfx || f.nfz != (f.z1 - f.z0 + fr - 1) / fr)
    return fail("region tiles", MX, MZ, T, f.x0, f.z0);
  // halos: stresses of level n+1 at [x0-6, x1+5] x [z0-6, z1+5] (interior
  // stress formulas: columns [3, MX-2], rows [3, MZ'-2]); velocities at
  // [x0-4, x1+3] x [z0-4, z1+3] (columns [4, MX-2], rows [2, MZ'-2])
  if (f.x0 - 6 < 3 || f.x1 + 5 > MX - 2 || f.z0 - 6 < 3 || f.z1 + 5 > MZ - 2)
    return fail("region stress halo", MX, MZ, T, f.x0, f.z0);
  if (f.x0 - 4 < 4 || f.x1 + 3 > MX - 2 || f.z0 - 4 < 2 || f.z1 + 3 > MZ - 2)
    return fail("region velocity halo", MX, MZ, T, f.x0, f.z0);
  // reads: velocities 8 cells around (rows to z1+6 for W and QZ of the last row)
  if (!inside(f.x0 - 8, f.z0 - 8) || !inside(f.x1 + 7, f.z1 + 7)) return fail("region reads", MX, MZ, T, f.x0, f.z0);
  for (int z = f.z0; z < f.z1; ++z)
    for (int x = f.x0; x < f.x1; ++x) { ++at(cs, x, z); ++at(cv, x, z); }
  long long nr_s = 0, nr_v = 0;
  for (int pass = 0; pass < 2; ++pass) {
    const Split &s = pass ? p.sv : p.ss;
    std::vector<int> &cnt = pass ? cv : cs;
    long long &n = pass ? nr_v : nr_s;
    const int zmin = pass ? 2 : 5;
    for (int q = 0; q < s.nrect; ++q) {
      const int x0 = s.qx0[q], x1 = s.qxl[q];
      if (x1 != x0 + 64 * s.qntx[q]) return fail("two-step run strips", MX, MZ, T, x0, s.qz0[q]);
      if ((x0 + off) % 32 != 0) return fail("run alignment", MX, MZ, T, x0, s.qz0[q]);
      if (x0 < 4 || x1 > MX - 1 || s.qz0[q] < zmin || s.qz1[q] > MZ - 1)
        return fail("run range", MX, MZ, T, x0, s.qz0[q]);
      for (int z = s.qz0[q]; z < s.qz1[q]; ++z)
        for (int x = x0; x < x1; ++x) { ++at(cnt, x, z); ++n; }
    }
    for (int r = 0; r < kMaxCells; ++r)
      for (int z = s.rz0[r]; z < s.rz1[r]; ++z)
        for (int x = s.rx0[r]; x < s.rx1[r]; ++x) {
          if (!inside(x, z)) return fail("cell outside grid", MX, MZ, T, x, z);
          ++at(cnt, x, z);
          ++n;
        }
  }
  if (nr_s != p.stress_cells || nr_v != p.velocity_cells) return fail("cell counts 2", MX, MZ, T, 0, 0);
  for (int z = 0; z <= g.PMLZH; ++z)
    for (int x = g.PMLL; x <= g.PMLXH; ++x) {
      const int want_s = z >= zs ? 1 : 0, want_v = (z >= zv && z <= g.PMLZH - 1) ? 1 : 0;
      if (at(cs, x, z) != want_s) return fail("stress coverage 2", MX, MZ, T, x, z);
      if (at(cv, x, z) != want_v) return fail("velocity coverage 2", MX, MZ, T, x, z);
    }
  return 0;
}

// plan_check pick MX MZ' T ...: the fused block height bq_create takes and
// whether the automatic choice keeps the fused tiles, one line per grid.
static int pick(int argc, char **argv) {
  for (int a = 2; a + 2 < argc; a += 3) {
    GridDesc g{};
    g.MX = std::atoi(argv[a]); g.MZ = std::atoi(argv[a + 1]); g.T = std::atoi(argv[a + 2]);
    g.PMLL = 1 - g.T; g.PMLXH = g.MX + g.T; g.PMLZH = g.MZ + g.T;
    g.NX = g.MX + 2 * g.T; g.NZ = g.MZ + g.T + 1;
    const Plan p = plan_cells(g, g.T - 1, 512, kTileRows[0], 4, 2, 2);
    std::printf("%d %d\n", fused_pick(p.fu), fused_worth(g, p.fu) ? 1 : 0);
  }
  return 0;
}

// plan_check est NFX ROWS FR: fused_estimate of nfx columns of full tiles
// over rows rows (no half or quarter tiles), in us.
static int est(char **argv) {
  Fused f{};
  f.nfx = std::atoi(argv[2]);
  f.vz0 = 0;
  f.vz1 = std::atoi(argv[3]);
  std::printf("%.6f\n", fused_estimate(f, std::atoi(argv[4])));
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 2 && std::string(argv[1]) == "pick") return pick(argc, argv);
  if (argc == 5 && std::string(argv[1]) == "est") return est(argv);
  if (argc < 9 || (argc - 6) % 3 != 0) {
    std::printf("usage: plan_check FW FR NR ZS ZV MX MZ T [MX MZ T ...]\n");
    return 2;
  }
  const int fw = std::atoi(argv[1]), fr = std::atoi(argv[2]), nr = std::atoi(argv[3]);
  const int zs = std::atoi(argv[4]), zv = std::atoi(argv[5]);
  int n = 0;
  for (int a = 6; a + 2 < argc; a += 3) {
    for (bool fused : {true, false})
      if (check(std::atoi(argv[a]), std::atoi(argv[a + 1]), std::atoi(argv[a + 2]), fw, fr, nr, zs, zv, fused))
        return 1;
    for (int w2 : {256, 192})
      if (check2(std::atoi(argv[a]), std::atoi(argv[a + 1]), std::atoi(argv[a + 2]), w2, fr, nr, zs, zv))
        return 1;
    ++n;
  }
  std::printf("ok %d\n", n);
  return 0;
}
