// bq_internal.h - shared definitions of the Biot time-loop library.
#pragma once
#include <cstdint>
#include <vector>

namespace bq {

// .MO table columns (read_model_par.F90:52-60).
enum {
  T_MU = 0, T_SXX1, T_SXX2, T_SXX3, T_SZZ1, T_SZZ2, T_SZZ3, T_P1, T_P2, T_P3,
  T_REL1_U, T_REL1_W, T_REL2_U, T_REL2_W, T_REL3_U, T_REL3_W,
  T_REL1_QX, T_REL1_QZ, T_REL2_QX, T_REL2_QZ, T_REL3_QX, T_REL3_QZ,
  T_AUXW, T_AUXQZ, T_NTAB
};

struct GridDesc {
  int MX, MZ, T, PMLL, PMLXH, PMLZH, NX, NZ, LIN_X, LIN_Z, KTBO, MT2, JMNUM;
  float H, DT;
};

// Host copies of the PML damping coefficients. x1: ODA/ODB/ODC/ODD _XL then
// _XH, each (0:T, NZc) as [l][j]; x2: same with 0:T+1; z1: ODA/ODB/ODC/ODD _ZH
// (0:T, NXc) as [i][j]; z2: same with 0:T+1.
struct PmlHost {
  std::vector<float> x1[8], x2[8], z1[4], z2[4];
  void resize(int T, int NX, int NZ) {
    for (auto &v : x1) v.assign((size_t)(T + 1) * NZ, 0.0f);
    for (auto &v : x2) v.assign((size_t)(T + 2) * NZ, 0.0f);
    for (auto &v : z1) v.assign((size_t)(T + 1) * NX, 0.0f);
    for (auto &v : z2) v.assign((size_t)(T + 2) * NX, 0.0f);
  }
};

// ---- partition of a step's cells among the kernels

// Cells [x0, x1) x [z0, z1).
struct Rect {
  int x0, x1, z0, z1;
};

constexpr int kMaxRuns = 4;   // run rectangles of one half-step
constexpr int kMaxCells = 8;  // per-cell rectangles of one half-step

// The cells of one half-step outside the fused tiles. Runs: up to kMaxRuns
// rectangles [qx0, qxl) x [qz0, qz1) of qntx 64-column strips from qx0 (the
// last strip ends past qxl when the width is not a multiple of 64: its lanes
// beyond qxl compute but do not store), in blocks of 64 columns x 4 runs of
// nr rows (qblk blocks, qcum cumulative). Per-cell path: up to kMaxCells
// rectangles [rx0, rx1) x [rz0, rz1), one thread per cell, cum the cumulative
// cell counts.
// Velocity half-step only (fold = 1): the edge rows come with it. etop top
// blocks take rows 2 and 3 and then row 1 (stress imaging), each over up to
// kEdgeCols columns: block k from column PMLL + kEdgeCols*k, the last one from
// elast to the grid's end, so that every column a block recomputes for its
// neighbour (one each side, row 2) uses the interior formulas. ebot bottom
// blocks take rows PMLZH-2 and PMLZH-1 and then row PMLZH with the bottom
// history, over kEdgeCols columns each.
constexpr int kEdgeCols = 128;

struct Split {
  int nr, nrect;
  int fold, etop, ebot, elast;
  int qx0[kMaxRuns], qxl[kMaxRuns], qntx[kMaxRuns], qz0[kMaxRuns], qz1[kMaxRuns], qblk[kMaxRuns];
  int qcum[kMaxRuns + 1];
  int rx0[kMaxCells], rx1[kMaxCells], rz0[kMaxCells], rz1[kMaxCells];
  int cum[kMaxCells + 1];
};

// Fused tiles over columns [vx0, vxe) and rows [vz0, vz1):
//   nfx columns of full tiles (the tile width fw) from vx0, in nfz rows of
//   blocks of the block height fr (the last one clipped);
//   then, where h2 = 1, one column of half tiles (fw/2 wide, fr/2 rows), two
//   to a block (nb2 blocks), and n4 (0..2) columns of quarter tiles (fw/4
//   wide, fr/4 rows), four to a block (nb4 blocks for all of them); the last
//   quarter column stores only its first w4 columns (w4 <= fw/4; its other
//   threads compute the halo values the stored columns need). A block's tiles
//   are consecutive in z; tiles past vz1 are empty.
// No fused columns: nfx = h2 = n4 = 0. The block counts depend on the block
// height (fused_rows).
struct Fused {
  int vx0, vxe, nfx, vz0, vz1, nfz;
  int h2, n4, w4, nb2, nb4;
};

// Sets nfz, nb2, nb4 of fu for block height fr.
void fused_rows(Fused &fu, int fr);

// Estimated duration (us) of a fused launch of block height fr: the blocks
// are dispatched in class order (full, half, quarter) into 512 slots (two
// blocks per CU), each marches its rows plus 4 halo rows one after another,
// and a row takes 1.9 us + 3.5 ns per running block (latency-bound with few
// blocks, bandwidth-bound with many; fitted on MI355X timings of grids from
// 3000 x 1500 to 8000 x 4000 cells at heights 16 to 128).
double fused_estimate(const Fused &fu, int fr);

// Fused block heights the planner and the autotuner choose from.
constexpr int kNumRows = 6;
constexpr int kTileRows[kNumRows] = {128, 64, 48, 32, 24, 16};

// The height of kTileRows whose launch fused_estimate finds shortest (the
// taller one on a tie).
int fused_pick(const Fused &fu);

// Grid cells (NXc x NZc) up to which the automatic choice leaves a grid
// without fused tiles. The runs and the per-cell path read and write both
// field sets once per half-step; while those (64 B per cell) stay within
// about the 256 MB last-level cache they outrun the fused tiles, whose
// blocks march their rows one after another and leave CUs idle on a small
// grid. Measured on MI355X: 4.33 M cells (3040 x 1423) step faster without
// fused tiles (39.9 against 39.3 Gcell/s), 4.63 M (3040 x 1523) with them
// (40.4 against 38.6).
constexpr long long kCellsMax = 4500000;

// Whether the automatic choice takes the fused tiles (of fu's columns, on
// grid g): where g holds more than kCellsMax cells and fu any fused block.
bool fused_worth(const GridDesc &g, const Fused &fu);

// Blocks of a fused launch: full, half and quarter tiles.
inline int fused_blocks(const Fused &fu) { return fu.nfx * fu.nfz + fu.nb2 + fu.nb4; }

// Blocks of a fused launch in XCD order (fused_item): each of the 8 XCDs
// takes its share of every class, so the launch is 8 times the largest share
// (the blocks past an XCD's share do nothing).
inline int fused_launch_blocks(const Fused &fu) {
  const int n[3] = {fu.nfx * fu.nfz, fu.nb2, fu.nb4};
  int l = 0;
  for (int c = 0; c < 3; ++c) l += (n[c] + 7) / 8;
  return 8 * l;
}

struct Plan {
  Fused fu;
  Split ss, sv;  // stress and velocity half-step outside the fused tiles
  long long fused_cells, stress_cells, velocity_cells;
};

// Partition of the cells of one step: the stress half-step covers x in
// [PMLL, PMLXH], rows [zs, PMLZH]; the velocity half-step rows [zv, PMLZH-1]
// (the rows outside these are the kernels' own edge rows). off = -PMLL; fw,
// fr: fused tile width and height; nr: rows per wavefront run. fold: the
// velocity half-step takes the edge rows too (Split::fold), where the grid
// allows it. fused: false leaves every cell to the runs and the per-cell
// path (no fused tiles).
Plan plan_cells(const GridDesc &g, int off, int fw, int fr, int nr, int zs, int zv, bool fold = true,
                bool fused = true);

// Two-step region: own cells [x0, x1) x [z0, z1) in nfx x nfz tiles of the
// tile width and fr rows (the last row of tiles clipped); nfx = 0: none.
struct Fused2 {
  int x0, x1, z0, z1, nfx, nfz, fr;
};

struct Plan2 {
  Fused2 f2;
  Split ss, sv;  // stress and velocity half-steps of every cell outside the region
  long long region_cells, stress_cells, velocity_cells;
};

// Partition for two steps per sweep: the region's cells are advanced twice by
// the two-step tiles of width fw2 and height fr2; every other cell (the band
// of interior cells around the region, the PML strips, the surface and bottom
// rows) by the per-cell and run kernels, one step at a time. The region keeps
// 6 cells of interior-formula cells on every side (its halos reach that far).
Plan2 plan_cells2(const GridDesc &g, int off, int fw2, int fr2, int nr, int zs, int zv);

// NONREF inputs from MODULE CONTROL_DATA: OMG as READ_OPEN leaves it (the
// namelist value / 2 pi, read_open.F90:188), WB, THPPLE, THSSLE.
struct NrParams {
  float OMG, WB, THPPLE, THSSLE;
};

// NONREF (nonref.F90:56-206): ax = AX01 AX02 AX10 AX11 AX12 AX20 AX21 AX22 of
// boundary key `key` for speed VEL (VELS, VELP: S and fast-P speeds). Returns
// 0, or -1 for an unknown key.
int nonref(int key, float H, float DT, float VEL, float VELS, float VELP, const NrParams &p,
           float ax[8]);

// Bottom part of PRE_NR_COEF1 (pre_nr_coef1.F90:185-201): ABW from the fast-P
// speed, ABH from the S speed of the material at (PMLL, MZ').
int bottom_nr(const GridDesc &g, const int32_t *jm, const float *tab, const NrParams &p,
              float abw[8], float abh[8]);

void wave_speeds(float sig1, float p, float ainv, float mu, float rel1, float rel2, float rel3,
                 float *vpf, float *vps, float *vs);
void build_pml(const GridDesc &g, const int32_t *jm, const float *tab, float r, float tau,
               float pw, float sf, float wc, PmlHost &out);

}  // namespace bq
