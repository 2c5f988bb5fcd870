/*
 * Oracle: CPU restatement of critic2's cube readers, TEST INFRASTRUCTURE
 * ONLY (the checker and CPU baseline of the GPU cube parser, never the
 * product path).
 *
 * oracle_read_cube follows grid3%read_cube (grid3mod@proc.f90:512-572):
 * two title records, "natoms x0" (:535), three "n(i) step(:,i)" records
 * (:543-548), natoms atom records and, for natoms < 0, one more record
 * (:549-552); then the list-directed read of n1*n2*n3 values in the order
 * (((f(i,j,k),k=1,n3),j=1,n2),i=1,n1) (:564).  Each value is converted with
 * the C library's strtod (correctly rounded), after mapping the Fortran
 * exponent letters D/d/Q/q to E; a bare signed exponent (1.0-3) is also
 * accepted, as Fortran reads it.
 * oracle_read_bincube follows grid3%read_bincube (:575-635) over gfortran's
 * unformatted sequential records (4-byte markers; negative leading markers
 * continue a record).
 *
 * Both return 0 on success, 1 on a read error; with f == NULL only the
 * header is returned.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "c2oracle.h"

static int is_blank(int c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

/* Reads a whole file into memory (NUL-terminated). */
static char *slurp(const char *path, size_t *len) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return NULL;
  fseeko(fp, 0, SEEK_END);
  size_t n = (size_t)ftello(fp);
  fseeko(fp, 0, SEEK_SET);
  char *b = malloc(n + 1);
  if (b && fread(b, 1, n, fp) != n) {
    free(b);
    b = NULL;
  }
  fclose(fp);
  if (b) b[n] = 0;
  *len = n;
  return b;
}

/* One Fortran real constant starting at s (length len) to double. */
static int conv_real(const char *s, size_t len, double *out) {
  char tmp[256];
  if (len == 0 || len >= sizeof tmp) return 1;
  memcpy(tmp, s, len);
  tmp[len] = 0;
  for (size_t k = 0; k < len; k++) {
    char c = tmp[k];
    if (c == 'D' || c == 'd' || c == 'Q' || c == 'q') tmp[k] = 'E';
  }
  /* bare signed exponent: a sign after a digit or '.' */
  for (size_t k = 1; k < len; k++) {
    if ((tmp[k] == '+' || tmp[k] == '-') && ((tmp[k - 1] >= '0' && tmp[k - 1] <= '9') || tmp[k - 1] == '.')) {
      if (len + 1 >= sizeof tmp) return 1;
      memmove(tmp + k + 1, tmp + k, len - k + 1);
      tmp[k] = 'E';
      len++;
      break;
    }
  }
  char *end;
  *out = strtod(tmp, &end);
  return *end != 0;
}

typedef struct {
  const char *b;
  size_t n, p;
} Rd;

static int skip_record(Rd *r) {
  if (r->p >= r->n) return 1;
  while (r->p < r->n && r->b[r->p] != '\n') r->p++;
  if (r->p < r->n) r->p++;
  return 0;
}

/* Next blank-delimited token from position q; returns its start/length. */
static int next_token(Rd *r, size_t *q, size_t *start, size_t *len) {
  while (*q < r->n && is_blank(r->b[*q])) (*q)++;
  if (*q >= r->n) return 1;
  *start = *q;
  while (*q < r->n && !is_blank(r->b[*q])) (*q)++;
  *len = *q - *start;
  return 0;
}

/* "integer real real real" from the next record on (list-directed). */
static int read_ir3(Rd *r, long *iv, double rv[3]) {
  size_t q = r->p, s, l;
  char tmp[64];
  if (next_token(r, &q, &s, &l) || l >= sizeof tmp) return 1;
  memcpy(tmp, r->b + s, l);
  tmp[l] = 0;
  char *end;
  *iv = strtol(tmp, &end, 10);
  if (*end) return 1;
  for (int a = 0; a < 3; a++)
    if (next_token(r, &q, &s, &l) || conv_real(r->b + s, l, &rv[a])) return 1;
  r->p = q;
  if (r->p < r->n) skip_record(r);
  return 0;
}

int oracle_read_cube(const char *path, int *nat, int *ismo, int n[3], double x0[3], double step[9], double *f) {
  size_t len;
  char *b = slurp(path, &len);
  if (!b) return 1;
  Rd r = {b, len, 0};
  int err = 1;
  long iv;
  double rv[3];
  if (skip_record(&r) || skip_record(&r)) goto done;
  if (read_ir3(&r, &iv, rv)) goto done;
  *ismo = iv < 0;
  *nat = (int)(iv < 0 ? -iv : iv);
  memcpy(x0, rv, sizeof rv);
  for (int i = 0; i < 3; i++) {
    if (read_ir3(&r, &iv, rv) || iv < 1) goto done;
    n[i] = (int)iv;
    memcpy(step + 3 * i, rv, sizeof rv);
  }
  for (int a = 0; a < *nat; a++)
    if (skip_record(&r)) goto done;
  if (*ismo && skip_record(&r)) goto done;
  if (!f) {
    err = 0;
    goto done;
  }
  {
    const int64_t n1 = n[0], n2 = n[1], n3 = n[2];
    size_t q = r.p, s, l;
    for (int64_t i = 0; i < n1; i++)
      for (int64_t j = 0; j < n2; j++)
        for (int64_t k = 0; k < n3; k++) {
          double v;
          if (next_token(&r, &q, &s, &l) || conv_real(b + s, l, &v)) goto done;
          f[i + n1 * (j + n2 * k)] = v;
        }
  }
  err = 0;
done:
  free(b);
  return err;
}

/* One logical record into dst (at most cap bytes kept); returns its length
 * or -1 at a read error. */
static int64_t read_record(FILE *fp, void *dst, size_t cap) {
  int64_t total = 0;
  for (;;) {
    int32_t lead, trail;
    if (fread(&lead, 4, 1, fp) != 1) return -1;
    int64_t l = lead < 0 ? -(int64_t)lead : lead;
    for (int64_t k = 0; k < l; k++) {
      int c = fgetc(fp);
      if (c == EOF) return -1;
      if ((size_t)(total + k) < cap) ((unsigned char *)dst)[total + k] = (unsigned char)c;
    }
    if (fread(&trail, 4, 1, fp) != 1) return -1;
    total += l;
    if (lead >= 0) return total;
  }
}

int oracle_read_bincube(const char *path, int *nat, int *ismo, int n[3], double x0[3], double step[9],
                        double *f) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return 1;
  int err = 1;
  unsigned char r1[28], r2[84], ra[36];
  int32_t na;
  if (read_record(fp, r1, sizeof r1) < 28) goto done;
  memcpy(&na, r1, 4);
  memcpy(x0, r1 + 4, 24);
  *ismo = na < 0;
  *nat = na < 0 ? -na : na;
  if (read_record(fp, r2, sizeof r2) < 84) goto done;
  memcpy(n, r2, 12);
  memcpy(step, r2 + 12, 72);
  for (int a = 0; a < *nat; a++)
    if (read_record(fp, ra, sizeof ra) < 36) goto done;
  if (f) {
    const size_t need = (size_t)n[0] * n[1] * n[2] * 8;
    if (read_record(fp, f, need) < (int64_t)need) goto done;
  }
  err = 0;
done:
  fclose(fp);
  return err;
}
