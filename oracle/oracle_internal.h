/* ORACLE — test infrastructure only. Internal declarations shared by the
 * oracle's translation units. */
#ifndef FNEC_ORACLE_INTERNAL_H
#define FNEC_ORACLE_INTERNAL_H

#include "fnec_oracle.h"

#define OR_MAX_FIELDS 16

#define OR_MAX_RAW 12
typedef struct {
    char mn[4];
    int nf;
    double f[OR_MAX_FIELDS];
    char raw[OR_MAX_RAW][40]; /* NT: the raw field text (card.rs NtCard::raw_fields) */
} or_card;

struct or_deck {
    int n, cap;
    or_card *cards;
};

uint32_t or_as_u32(double v);
int or_parse_f64_strict(const char *s, double *out);
int32_t or_as_i32(double v);
void or_deck_rp_flags(const or_deck *d, int *normalize, int *avg_gain);

/* Physical constants exactly as the reference spells them. */
#define OR_PI 3.14159265358979323846
#define OR_C0 299792458.0
#define OR_MU0 (4.0 * OR_PI * 1e-7)
#define OR_EPS0_MATRIX 8.854187817e-12   /* matrix.rs:33 */
#define OR_EPS0_FF 8.8541878128e-12      /* farfield.rs:293 */
#define OR_ETA0 (OR_MU0 * OR_C0)         /* excitation.rs:21, farfield.rs:295 */

/* EX card accessors (nec_model/src/card.rs:76-107). */
static inline int ex_type(const or_card *c) { return (int)or_as_u32(c->f[0]); }
static inline uint32_t ex_tag(const or_card *c) { return or_as_u32(c->f[1]); }
static inline uint32_t ex_seg(const or_card *c) { return or_as_u32(c->f[2]); }
static inline double ex_vr(const or_card *c) { return c->nf > 4 ? c->f[4] : 0.0; }
static inline double ex_vi(const or_card *c) { return c->nf > 5 ? c->f[5] : 0.0; }
static inline int ex_is_voltage(const or_card *c) { int t = ex_type(c); return t == 0 || t == 5; }
static inline int ex_is_plane_wave(const or_card *c) { int t = ex_type(c); return t >= 1 && t <= 3; }
static inline int ex_is_current(const or_card *c) { return ex_type(c) == 4; }

int or_find_seg(const or_seg *s, int n, uint32_t tag, uint32_t idx);

extern int or_threads; /* row-loop threads (linear.c) */

#endif
