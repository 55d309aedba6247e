/*
 * Launch-shape choice for the ELLPACK row-product kernels.
 */
#ifndef BML_HIP_TUNE_H
#define BML_HIP_TUNE_H

/* Pick lanes per row (group) and LDS hash slots per row (table) for a
 * product whose left rows hold up to max_a entries and right rows up to
 * max_b; touched_hint is the largest distinct-column count seen the last
 * time (0 = unknown). */
void bmlh_pick_spgemm(int precision, int N, int max_a, int max_b, int touched_hint,
                      int *group, int *table);

#endif
