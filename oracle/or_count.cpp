/* or_count.cpp -- the counting build of the oracle (TEST INFRASTRUCTURE ONLY):
 * elegant_oracle.c compiled as C++ with counted doubles (or_count.h), its
 * entry points kept with C linkage so that oracle.py loads it like the other
 * two builds. */
#include "or_count.h"

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

#include "elegant_oracle.c"
}
