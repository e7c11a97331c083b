/*
 * biotfd.h - C ABI of the MI355X time loop for the 2-D Biot velocity-stress
 * simulator (drop-in for SUBROUTINE TIME_LOOP, time_loop.F90:1-209).
 *
 * The reference's driver calls TIME_LOOP with no arguments (main.F90:84;
 * interface at mod_interfaces.F90:1265) and keeps all state in Fortran modules.
 * This ABI replaces the body of the step loop (time_loop.F90:91-190) with a
 * handle that owns the device-resident grid. A Fortran TIME_LOOP replacement
 * binds these entry points with ISO_C_BINDING (scalars by VALUE, arrays as
 * assumed-size dummies); the interface block is MODULE BQ_BINDING in
 * 2d-vqsp-fd-sim_amd/fortran/bq_binding.f90.
 *
 * Conventions
 *  - Fortran (x,z) arrays are C [z][x] arrays with the lower bounds removed:
 *    element (i,l) of a full-grid array sits at l*NXc + (i-PMLL), with
 *    PMLL = 1-TPML, NXc = MX+2*TPML, NZc = MZ'+TPML+1, MZ' as after
 *    read_open.F90:149-151 (MZ+2 with stress imaging).
 *  - Every real is IEEE binary32 (the reference's PP = SP, mod_precision.F90).
 *  - Calls return 0 on success and a negative code on failure; the message is
 *    available from bq_last_error. The reference reports errors by writing to
 *    its log unit and STOPping (tl_alloc.F90:39-42); the Fortran binding does
 *    the same with this message.
 *  - One handle owns its device memory and one HIP stream on the device that
 *    was current at bq_create. Calls on one handle must not overlap; distinct
 *    handles may run concurrently.
 */
#ifndef BIOTFD_H
#define BIOTFD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bq_handle bq_handle;

/* Run description. Fields follow MODULE CONTROL_DATA (mod_contr_data.F90:13-31)
 * and MODULE AUXIL (mod_auxil.F90:24-33) after READ_OPEN and AUX1. */
typedef struct bq_desc {
  int32_t mx;             /* MX: physical cells in x */
  int32_t mz;             /* MZ' : MZ after read_open.F90:149-151 */
  int32_t tpml;           /* TPML (1..256) */
  int32_t stress_imaging; /* STRESS_IMAGING (logical as 0/1) */
  int32_t lin_x, lin_z;   /* source cell, read_open.F90:163-169 */
  int32_t ktbo;           /* KTBO: bottom boundary key (0 = PML only) */
  int32_t mt2;            /* MT2: number of time steps */
  int32_t jmnum;          /* JMNUM: entries in the material table */
  int32_t variant;        /* BQ_AFDA_* bits: the reference's compile-time options */
  double h, dt;           /* H, DT (binary32 values widened to double) */
} bq_desc;

/* bq_desc.variant: the reference's build options. BQ_AFDA_ZZ, BQ_AFDA_XZ: the
 * one-sided stencils of the explicit free surface (stress_imaging = 0),
 * -DAFDA_ZZ / -DAFDA_XZ: STRAIN_UW_ZZ_0 / STRAIN_Q_ZZ_0 (strain_uw_zz_0.F90:52-61,
 * strain_q_zz_0.F90:47-56) and STRAIN_UW_XZ_1 (strain_uw_xz_1.F90:83-96).
 * BQ_SECOND: the second-order build -DSECOND, A = 0, B = 1 (mod_auxil.F90:15-21)
 * with the 2-point row-0 forms of inh_qx_int_0.F90:40-47 and
 * inh_v_uv_int_0.F90:40-48. */
enum { BQ_AFDA_ZZ = 1, BQ_AFDA_XZ = 2, BQ_SECOND = 4 };

/* Field ids for bq_get_field / bq_set_field (mod_wavefield.F90:20,
 * mod_stressfield.F90:8). */
enum {
  BQ_U = 0, BQ_W = 1, BQ_QX = 2, BQ_QZ = 3,
  BQ_TXX = 4, BQ_TZZ = 5, BQ_TXZ = 6, BQ_PRES = 7
};

/* Allocates and zeroes the grid (TL_ALLOC, tl_alloc.F90:24-56; PML memories
 * as ppml.F90:115-153; bottom history as alloc_nr.F90:41-52). */
int bq_create(const bq_desc *desc, bq_handle **out);

/* READ_MODEL + READ_MODEL_PAR results (read_model.F90:57-87,
 * read_model_par.F90:37-74): jm is int32 [NZc][NXc] (1-based material ids,
 * already replicated into PML and image rows), tab is float [JMNUM][24] in the
 * .MO order, with REL* entries already scaled by DT/H or DT. The friction
 * factors EXP(-REL3_Q*) and (REL2/REL3_Q)*(1-EXP(-REL3_Q)) are formed here on
 * the host in binary32, in the reference's expression order
 * (inh_qx_int.F90:44, inh_v_uv_int.F90:46). */
int bq_upload_model(bq_handle *h, const int32_t *jm, const float *tab);

/* PPML (ppml.F90:159-306): builds the 24 damping-coefficient arrays from the
 * uploaded model, in binary32, and uploads them. Needs bq_upload_model first. */
int bq_build_pml(bq_handle *h, float r, float tau, float pow_, float sf, float wc);

/* Alternative to bq_build_pml: the 24 arrays as the Fortran host computed
 * them, in MODULE PML order (mod_pml.F90:7-19): ODA_XL ODB_XL ODC_XL ODA_XH
 * ODB_XH ODC_XH ODD_XL ODD_XH (each (0:T, NZc)), ODA_ZH ODB_ZH ODC_ZH ODD_ZH
 * ((0:T, NXc)), ODA2_XL ODB2_XL ODC2_XL ODA2_XH ODB2_XH ODC2_XH ODD2_XL ODD2_XH
 * ((0:T+1, NZc)), ODA2_ZH ODB2_ZH ODC2_ZH ODD2_ZH ((0:T+1, NXc)). */
int bq_upload_pml(bq_handle *h, const float *const *coef);

/* Copies one of the 24 PML coefficient arrays back (same order); returns the
 * element count, or a negative code. out may be NULL to query the size. */
int bq_get_pml(bq_handle *h, int id, float *out);

/* Bottom non-reflecting coefficients ABW01..ABW22 and ABH01..ABH22
 * (mod_nr_bound.F90:47-52, PRE_NR_COEF1 pre_nr_coef1.F90:185-201), each in the
 * order 01 02 10 11 12 20 21 22. All zero for KTBO = 0. */
int bq_upload_bottom_nr(bq_handle *h, const float abw[8], const float abh[8]);

/* Alternative to bq_upload_bottom_nr: PRE_NR_COEF1's bottom part
 * (pre_nr_coef1.F90:185-201, NONREF nonref.F90:56-206 with HIGDON_FD,
 * HIGCE_FD, HIGCE2_FD, OptAbsBc) computed on the host from the uploaded model
 * for key desc.ktbo (0..9), in the reference's binary32 / binary64 order, and
 * uploaded. prm = OMG as READ_OPEN leaves it (namelist OMG / 2 pi,
 * read_open.F90:188), WB, THPPLE, THSSLE (read_open.F90:173-179). abw, abh
 * receive the coefficients if not NULL. Needs bq_upload_model first. */
int bq_build_bottom_nr(bq_handle *h, const float prm[4], float abw[8], float abh[8]);

/* SOURTF_A(2,MT2) (open_sr_file.F90:78-110): per step, solid then fluid. */
int bq_set_source(bq_handle *h, const float *sourtf);

/* Receivers (read_open.F90:234-236): 1-based x index and z index of the U
 * node (IRECX, LRECX) and of the W node (IRECZ, LRECZ). Allocates a device
 * trace buffer for MT2 steps. Part of the set-up: rejected (-6) once a step
 * has run; a second call before the first step replaces the receivers and
 * frees the previous buffers. */
int bq_set_receivers(bq_handle *h, int mr, const int32_t *irecx, const int32_t *lrecx,
                     const int32_t *irecz, const int32_t *lrecz);

/* Advances steps itile_begin..itile_end (1-based, DO ITILE of
 * time_loop.F90:65); steps must follow one another. Returns after queuing;
 * bq_sync waits. */
int bq_run(bq_handle *h, int itile_begin, int itile_end);
int bq_sync(bq_handle *h);

/* Receiver samples of steps first_step .. first_step+nsteps-1, [step][mr][6]:
 * UM(IRECX,LRECX), -WM(IRECZ,LRECZ), QX(IRECX,LRECX), -QZ(IRECZ,LRECZ),
 * PRES(IRECZ,LRECX), -TXX(IRECZ,LRECX) (store_seismograms.F90:58-83,
 * store_seismograms_pressure.F90:50-56). Out-of-grid receivers read 0. */
int bq_get_traces(bq_handle *h, float *out, int first_step, int nsteps);

/* Whole-grid copy of one field, [NZc][NXc] (snapshots, tests, restart). */
int bq_get_field(bq_handle *h, int field_id, float *host);
int bq_set_field(bq_handle *h, int field_id, const float *host);

/* Host-side set-up helpers (no handle, no device work). */
/* OPEN_SR_FILE (open_sr_file.F90:41-116): clamped cubic spline of the nstf
 * STF.DAT samples onto t=(n-1)*DT, scaled by KEY_SOUR/QR/H, -> sourtf[2*mt2]. */
int bq_source_from_stf(const float *t, const float *v, int nstf, int mt2, float dt, float h,
                       int key_sour, float qr, float *sourtf);

/* NONREF (nonref.F90:56-206) for one boundary and key 0..9: out = AX01 AX02
 * AX10 AX11 AX12 AX20 AX21 AX22 for speed vel (vels, velp: the S and fast-P
 * speeds of the cell), prm as for bq_build_bottom_nr. */
int bq_nonref(int key, float h, float dt, float vel, float vels, float velp, const float prm[4],
              float out[8]);

/* ---- Batches: many shots advanced together ------------------------------
 *
 * The reference runs one shot per process and splits one grid over MPI ranks
 * (domain_decomposition.F90:40-143); it has no batching. A batch groups the
 * handles of independent shots - each a whole TIME_LOOP (time_loop.F90:65-206)
 * of its own model, grid size, source and receivers - on one device and
 * advances all of them with three launches per time step in total, replayed
 * from a HIP graph. Every shot's fields and traces are bit-identical to a run
 * of that shot alone with bq_run. The handles stay the shots' own: their
 * results are read with bq_get_field / bq_get_traces as usual, and after a
 * batch run they continue at their next step. While a batch runs, its shots
 * must not be used otherwise. */
typedef struct bq_batch bq_batch;

/* Schedules of the batched launches: one block per work item, or a persistent
 * grid of a few blocks per CU taking work items from per-XCD queues. */
enum { BQ_SCHED_GRID = 0, BQ_SCHED_PERSISTENT = 1 };

/* Groups nshots distinct handles of one device (each with model, PML and
 * source set, receivers optional). The batch keeps the pointers; destroy it
 * before its handles. */
int bq_batch_create(bq_handle *const *shots, int nshots, bq_batch **out);

/* Advances every shot by nsteps steps from its own next step (the shots may
 * be at different steps; none may pass its MT2). Queued on the batch's stream;
 * later calls on the shots' handles are ordered after it. */
int bq_batch_run(bq_batch *b, int nsteps);
int bq_batch_sync(bq_batch *b);

/* Device time of the last bq_batch_run, ms (waits for it). */
int bq_batch_last_run_ms(bq_batch *b, double *ms);

/* Autotuner: runs trial_steps real steps of the batch with each of the six
 * candidates - fused tile height 64, 128, 256 rows x schedule grid, persistent
 * (candidate c: height index c / 2, schedule c % 2) - times each with HIP
 * events and keeps the fastest. The 6*trial_steps steps advance the shots like
 * any others (every candidate gives identical results). ms_per_step[6], if not
 * NULL, receives each candidate's device time per step. */
int bq_batch_autotune(bq_batch *b, int trial_steps, double *ms_per_step);

/* The schedule by hand: tile_rows 64, 128 or 256, schedule BQ_SCHED_*,
 * use_graph 0/1 (replay steps from a HIP graph). */
int bq_batch_set_schedule(bq_batch *b, int tile_rows, int schedule, int use_graph);
int bq_batch_get_schedule(bq_batch *b, int *tile_rows, int *schedule, int *use_graph);

/* Message of the last failure; NULL for a failed bq_batch_create. */
const char *bq_batch_last_error(bq_batch *b);
void bq_batch_destroy(bq_batch *b);

/* Misfit of two runs' receiver traces over steps first_step ..
 * first_step+nsteps-1, computed on the device where the traces lie (no
 * counterpart in the reference, which only writes its seismograms to files):
 * total[0] = sum over those steps, the receivers and the six columns of
 * bq_get_traces of (a - b)^2, total[1] = sum of b^2, accumulated in binary64;
 * per_receiver[2j], [2j+1] (NULL: not returned) the same for receiver j.
 * sqrt(total[0]/total[1]) is the relative L2 misfit of a against b. Both runs
 * must have the same receiver count and have run those steps; b's queued work
 * is waited for. Allocates a staging buffer of 14*mr+2 doubles per call. */
int bq_trace_misfit(bq_handle *a, bq_handle *b, int first_step, int nsteps, double *per_receiver,
                    double total[2]);

/* bq_trace_misfit without any allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops). Stages in a fixed device array of 917,520
 * bytes per device (held until the process ends), so it takes at most 8192
 * receivers (-3 beyond); calls that overlap on one device race on that
 * array. Same results as bq_trace_misfit. */
int bq_trace_misfit_noalloc(bq_handle *a, bq_handle *b, int first_step, int nsteps,
                            double *per_receiver, double total[2]);

/* Output writers (host side, no handle). STORE_SEISMOGRAMS and
 * STORE_SEISMOGRAMS_PRESSURE (store_seismograms.F90:47-154,
 * store_seismograms_pressure.F90:45-83) for a chunk of steps: samples is
 * [nsteps][mr][6] as bq_get_traces returns it for steps first_step ..
 * first_step+nsteps-1; appends one line per step to REC_UW_<j>.asc (t, U, -W
 * and their time derivatives), REC_Q_<j>.asc with key_tlq (QX, -QZ) and
 * REC_STRESS_<j>.asc with key_tlp (PRES, -TXX) in directory dir, in the
 * reference's "(n(1X,E13.6))" format, opening each file once per call.
 * inside: [mr][2] int32, whether the U node and the W node lie in the grid
 * (NULL: all); outside nodes give 0 samples and derivatives. state: [mr][4]
 * binary32 SSU, SSW, SSQX, SSQZ (the previous samples the derivatives are
 * formed from), zero before the first step, updated in place. Receivers are
 * written by up to 16 host threads. */
int bq_write_seismograms(const char *dir, const float *samples, int first_step, int nsteps,
                         int mr, float dt, const int32_t *inside, int key_tlq, int key_tlp,
                         float *state);
/* The Fortran E13.6 edit of one binary32 value as the reference's runtime
 * writes it (13 characters and a terminating NUL into out[14]). */
int bq_format_e13_6(float x, char *out);

/* Informational. */
const char *bq_last_error(bq_handle *h);
int bq_grid_shape(bq_handle *h, int32_t *nx, int32_t *nz);
/* Average device time per step of each launch of the last bq_run, ms, from
 * HIP events on the handle's stream: out[0] k_step (both half-steps of the
 * interior cells, the stress half-step of all other cells and the previous
 * step's receiver samples; on a partition without fused tiles, bq_set_tiles,
 * k_rest_stress: the stress half-step of every cell and the samples), out[1]
 * k_velocity (velocity half-step of the other
 * cells, with rows 1 and PMLZH where its edge blocks take them), out[2]
 * k_edges (rows 1 and PMLZH; 0 where k_velocity takes them: every grid with
 * TPML <= 124 and at least 3 rows between the surface and bottom edge blocks,
 * in single steps). Timing must have been enabled
 * with bq_enable_timing before that run: every = n >= 1 brackets the launches
 * of every n-th step of a run (its first step included) with events, 0 turns
 * timing off. */
int bq_enable_timing(bq_handle *h, int every);
int bq_kernel_times(bq_handle *h, double out[3]);
/* Cells updated per step: out[0] by both half-steps inside k_step (fused
 * blocks; 0 without fused tiles), out[1] by the stress half-step of k_step
 * outside them (of k_rest_stress without fused tiles), out[2] by k_velocity. */
int bq_kernel_cells(bq_handle *h, long long out[3]);

/* Two steps per sweep (no counterpart in the reference, whose loop advances
 * one step at a time, time_loop.F90:65-206). Where a grid is large enough, the
 * interior cells at least 6 cells inside the interior-formula region (columns
 * [9, MX-7), rows [9, MZ'-7)) are advanced two steps per pass over memory;
 * the other cells one step at a time, as always. Results are bit-identical
 * to single steps. bq_run uses it for each pair of steps after
 * bq_set_sweep(h, 2) when the handle allows it: no learned boundary attached,
 * and every receiver node outside that region or within 2 cells of its edge.
 * The default, 1, keeps single steps (on MI355X they are faster: the
 * two-step pass moves half the bytes per step but is bound by the latency of
 * its longer chain per row). bq_get_sweep returns the steps per sweep
 * bq_run will use (1 or 2). */
int bq_set_sweep(bq_handle *h, int steps);
int bq_get_sweep(bq_handle *h);
/* Fused tiles or not (no counterpart in the reference, whose loops update every
 * cell alike, time_loop.F90:65-206). Each step's interior is either marched in
 * fused tiles (both half-steps per pass, k_step) or left to the runs and the
 * per-cell path of the two half-steps; results are bit-identical either way.
 * BQ_TILES_AUTO (the default) takes the fused tiles on grids of more than
 * 4.5 M cells (NXc x NZc, about 3000 x 1500 and up): on a smaller grid a
 * fused launch lasts as long as one block's march down its rows while CUs
 * idle, and both field sets still fit about the last-level cache. BQ_TILES_FUSED and BQ_TILES_CELLS force one or the
 * other. bq_create picks the fused blocks' height (128, 64, 48, 32, 24 or 16
 * rows) from a model of the launch. The mode may be changed between runs. Batches
 * (bq_batch_create) always use the fused tiles. bq_get_tiles returns
 * BQ_TILES_FUSED or BQ_TILES_CELLS for the partition bq_run uses
 * (BQ_TILES_CELLS where the grid has no fused tiles). */
enum { BQ_TILES_AUTO = 0, BQ_TILES_FUSED = 1, BQ_TILES_CELLS = 2 };
int bq_set_tiles(bq_handle *h, int mode);
int bq_get_tiles(bq_handle *h);
/* Rows per fused block of the partition with fused tiles (128, 64, 48, 32, 24
 * or 16: bq_create's choice, or bq_autotune's), whether or not bq_run uses
 * that partition. */
int bq_get_tile_rows(bq_handle *h);
/* Autotuner (no counterpart in the reference): runs 1 + trial_steps real
 * steps with each of seven partitions - fused blocks of 128, 64, 48, 32, 24
 * and 16 rows (candidates 0-5), then no fused tiles (candidate 6) - times the
 * last trial_steps of each with HIP events and keeps the fastest
 * (bq_set_tiles(h, BQ_TILES_FUSED or BQ_TILES_CELLS) with the fastest
 * height). The 7 * (trial_steps + 1) steps advance the run like any others
 * (every candidate gives identical results); they are single steps whatever
 * bq_set_sweep allows. ms_per_step[7], if not NULL, receives each candidate's
 * device time per step. If a trial fails, the partition stays as it was and
 * the steps already run stay run (bq_run's next step follows them). */
int bq_autotune(bq_handle *h, int trial_steps, double *ms_per_step);
/* Timing of the last timed bq_run by launch kind, ms per launch. With two
 * steps per sweep (*steps_per_sweep = 2): ms[0] the two-step sweep of the
 * region (two steps), ms[1] the stress half-step of the other cells with the
 * samples, ms[2] their velocity half-step, ms[3] the edge rows (the last three
 * once per step); cells[0..2] = region cells, other stress cells, other
 * velocity cells. With single steps: ms[0..2] as bq_kernel_times, ms[3] = 0,
 * cells[0..2] as bq_kernel_cells. cells[3] = 0. bq_kernel_times returns
 * ms[0], ms[2], ms[3] of a two-step run. */
int bq_kernel_stats(bq_handle *h, double ms[4], long long cells[4], int *steps_per_sweep);
void bq_destroy(bq_handle *h);

#ifdef __cplusplus
}
#endif
#endif /* BIOTFD_H */
