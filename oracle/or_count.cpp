/* or_count.cpp -- the counting build of the oracle (TEST INFRASTRUCTURE ONLY):
 * elegant_oracle.c compiled as C++ with counted doubles (or_count.h), its
 * entry points kept with C linkage so that oracle.py loads it like the other
 * two builds. */
#include "or_count.h"
#undef double /* the counters and the argument log hold plain doubles */

extern "C" {
uint64_t or_cnt[6], or_cnt_all[6];
int or_cnt_matrix;

void or_count_reset(void) {
  memset(or_cnt, 0, sizeof or_cnt);
  memset(or_cnt_all, 0, sizeof or_cnt_all);
}

/* out[0..5]: add, sub, mul, div, sqrt, transcendental of the census; out[6..11]:
 * the same, every operation executed */
void or_count_read(uint64_t *out) {
  memcpy(out, or_cnt, sizeof or_cnt);
  memcpy(out + 6, or_cnt_all, sizeof or_cnt_all);
}

double *or_targ;
int32_t *or_tkind;
long or_targ_n, or_targ_cap;

/* log the arguments of up to cap transcendental calls (cap 0: stop logging) */
void or_targ_start(long cap) {
  free(or_targ);
  free(or_tkind);
  or_targ = cap > 0 ? (double *)malloc(sizeof(double) * cap) : NULL;
  or_tkind = cap > 0 ? (int32_t *)malloc(sizeof(int32_t) * cap) : NULL;
  or_targ_cap = or_targ && or_tkind ? cap : 0;
  or_targ_n = 0;
}

long or_targ_read(double *args, int32_t *kinds, long n) {
  long k = n < or_targ_n ? n : or_targ_n;
  memcpy(args, or_targ, sizeof(double) * k);
  memcpy(kinds, or_tkind, sizeof(int32_t) * k);
  return or_targ_n;
}

#define double or_cd
#include "elegant_oracle.c"
}
