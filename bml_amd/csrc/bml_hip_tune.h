/*
 * Launch-shape choice for the ELLPACK row-product kernels.
 */
#ifndef BML_HIP_TUNE_H
#define BML_HIP_TUNE_H

typedef struct bmlh_spgemm_shape {
    int group;  /* lanes per output row */
    int table;  /* LDS hash slots per row */
    int bucket; /* slots per probe */
    int remap;  /* XCD-aware block order */
    int waves;  /* wavefronts per row (long rows; unordered engine) */
    int engine; /* BMLHIP_ROW_EXACT, _UNORDERED or _WINDOW */
    int wbits;  /* _WINDOW: log2 of the column window */
    int wrows;  /* _WINDOW: output rows per block */
} bmlh_spgemm_shape;

/* Launch shape for a product whose left rows hold up to max_a entries and
 * right rows up to max_b; touched_hint is the largest distinct-column count
 * seen the last time (0 = unknown), when the left rows held up to
 * hint_max_a entries.  The engine follows bml_hip_set_row_engine unless
 * flags holds BMLH_PICK_EXACT; BMLH_PICK_SQUARE marks a product X*X (x2, the
 * SP2 iterations), the only one the window engine takes. */
#define BMLH_PICK_SQUARE 1
#define BMLH_PICK_EXACT 2
void bmlh_pick_spgemm(int precision, int N, int max_a, int max_b, int touched_hint,
                      int hint_max_a, int flags, bmlh_spgemm_shape *out);

/* Runs the product with shape sh on the caller's operands (its result
 * overwrites the output, so the caller runs the chosen shape last) and
 * gives its duration in ms; nonzero if the shape does not launch. */
typedef int (*bmlh_measure_fn)(const bmlh_spgemm_shape *sh, void *ctx, double *ms);

/* bmlh_pick_spgemm, then, with autotuning on (bml_hip_set_autotune or
 * BML_HIP_AUTOTUNE=1) and no shape set through the environment, the
 * measured shape for this key: from the cache, or searched with fn. */
void bmlh_tuned_spgemm(int precision, int N, int max_a, int max_b, int touched_hint,
                       int hint_max_a, int flags, bmlh_spgemm_shape *out, bmlh_measure_fn fn,
                       void *ctx);

#endif
