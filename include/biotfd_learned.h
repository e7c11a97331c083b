/*
 * biotfd_learned.h - learned absorbing boundary (extension of the time-loop
 * library, libbiotfd.so). No counterpart in the reference, which absorbs with
 * a CPML strip of TPML cells (ppml.F90:159-306, used by every HPLANE_*_PML
 * row kernel).
 *
 * A grid created with TPML = 8 is the reference's method with an 8-cell CPML.
 * With a network attached, its 8 strip cells per side (left x = -7..0, right
 * x = MX+1..MX+8, bottom rows MZ'+1..MZ'+8, corners with the bottom) are
 * replaced after every step by the network's prediction, made from the last
 * four time levels of a 12-cell band across each boundary (the strip and the
 * 4 interior cells next to it, all 8 fields). The networks - one per side,
 * three 5-tap convolutions along the boundary, 384 -> 64 -> 64 -> 64 channels,
 * leaky ReLU, no bias - are trained on bands recorded from runs of the same
 * models with a wide PML; the trainer below does that on the device (FP32
 * forward, backward and Adam kernels). Inference runs in BF16 on the matrix
 * cores with FP32 accumulation.
 *
 * Band of one time level (bq_band_size floats): left side [MZ'+1][8][12],
 * right side [MZ'+1][8][12], bottom [MX+16][8][12]; [position along the
 * boundary][field U W QX QZ TXX TZZ TXZ PRES][cell across it, 0 = outermost
 * strip cell .. 7 = innermost, 8..11 = interior cells inward].
 *
 * Weights (bq_lb_create, trainer): per side (left, right, bottom) W1
 * [5][384][64], W2 [5][64][64], W3 [5][64][64], FP32, [tap][in][out] with out
 * fastest; tap t multiplies the input at position p + t - 2. Input channel
 * level*96 + field*12 + cell (level 0 = oldest), output channel field*8 +
 * cell. Scales: [3][8], one per side and field; inputs are divided by them and
 * the outputs multiplied. Three forms: direct (0: the output times gain is
 * the strip's change over the step), residual (1: the output times gain
 * corrects the value the 8-cell CPML has just computed for the strip cell;
 * gain 0 is the plain CPML) and damping (2: the residual correction limited
 * to the interval between 0 and the CPML's value, so the network can only
 * damp the strip, like a sponge whose per-cell factors it picks).
 */
#ifndef BIOTFD_LEARNED_H
#define BIOTFD_LEARNED_H

#include <stdint.h>

#include "biotfd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bq_lb bq_lb;
typedef struct bq_lb_trainer bq_lb_trainer;

/* Floats in one time level of the band of a grid with MX, MZ'. */
int bq_band_size(int mx, int mzp);

/* The band of the grid's current fields (TPML >= 8) into device memory dst,
 * queued on the handle's stream. */
int bq_band_extract(bq_handle *h, float *dst);

/* Networks of the three sides from FP32 weights and scales (layout above),
 * converted to BF16 on the device that is current; form 0/1/2 (direct,
 * residual, damping), gain (0..16) scales the output, inner (1..8) is the
 * number of strip cells, counted from the interior, the output is applied to
 * (the others keep their update). */
int bq_lb_create(const float *weights, const float *scales, int form, float gain, int inner,
                 bq_lb **out);
void bq_lb_destroy(bq_lb *lb);
const char *bq_lb_last_error(void);

/* Attaches the networks to a grid of TPML = 8 before its first step: from
 * then on bq_run and bq_batch_run replace the strip cells after each step.
 * bq_lb_detach returns the grid to its plain 8-cell CPML. */
int bq_lb_attach(bq_handle *h, bq_lb *lb);
int bq_lb_detach(bq_handle *h);

/* The inference kernel on given bands of one side (diagnostics and tests):
 * bands [4][len][96] (oldest level first), out [len][8][8] (field, cell) the
 * newest level's strip values plus gain * scale * the network's output. */
int bq_lb_infer(bq_lb *lb, int side, int len, const float *bands, float *out);

/* The 8-cell strip (all fields) of dst's current fields from src's, at the
 * same physical cells (same MX, MZ'; both TPML >= 8). Used to keep a
 * truncated grid on the trajectory of a wide-PML run while recording. */
int bq_strip_copy(bq_handle *dst, bq_handle *src);

/* Misfit of one field (BQ_U ..) between two grids of the same physical size
 * over their physical cells x = 1..MX, rows 0..MZ' (their PML widths may
 * differ): out[0] = sum (a-b)^2, out[1] = sum b^2. */
int bq_field_misfit(bq_handle *a, bq_handle *b, int field_id, double out[2]);

/* ---- trainer: bands of wide-PML runs, fitted on the device ---------------
 * Groups of five levels of one run; levels 0..3 (t-3 .. t) are the input.
 * Direct form: level 4 is level t+1 and the target the strip's change from
 * level 3 to 4. Residual form: level 4 is the wide run's band at t+1 minus
 * the band of an 8-cell grid kept on its trajectory (bq_strip_copy) after
 * that grid's own step t+1 (bq_lb_trainer_capture_residual), and the target
 * its strip cells. All runs recorded must have the trainer's MX, MZ' and
 * TPML >= 8. inner (1..8): only the targets of the strip cells nearest the
 * interior enter the loss. */
int bq_lb_trainer_create(int mx, int mzp, int max_groups, int batch, int residual, int inner,
                         bq_lb_trainer **out);
void bq_lb_trainer_destroy(bq_lb_trainer *t);
const char *bq_lb_trainer_last_error(bq_lb_trainer *t);
/* Band of h's current fields as level `level` (0..4) of group `group`. */
int bq_lb_trainer_capture(bq_lb_trainer *t, bq_handle *h, int group, int level);
/* Level 4 of group `group`: band(wide) - band(trunc), both current fields. */
int bq_lb_trainer_capture_residual(bq_lb_trainer *t, bq_handle *wide, bq_handle *trunc, int group);
int bq_lb_trainer_sync(bq_lb_trainer *t);
/* RMS per side and field over levels 0..3 of groups 0..ngroups-1: the scales
 * (written to scales[3][8] and used from then on). */
int bq_lb_trainer_scales(bq_lb_trainer *t, int ngroups, float *scales);
int bq_lb_trainer_set_weights(bq_lb_trainer *t, const float *weights);
int bq_lb_trainer_get_weights(bq_lb_trainer *t, float *weights);
/* One minibatch of `batch` groups (indices) of side `side`: mean squared error
 * of the scaled target over the inner cells into *loss; with lr > 0 also backward and an Adam step
 * (beta 0.9, 0.999). */
int bq_lb_trainer_step(bq_lb_trainer *t, int side, const int32_t *groups, float lr, double *loss);

#ifdef __cplusplus
}
#endif
#endif /* BIOTFD_LEARNED_H */
